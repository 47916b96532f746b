// Engine: many simulations on one GPU (kw_engine_* / kw_submit / kw_wait in
// include/kwhip.h).
//
// Admission threads take submitted problems in order and move them to the
// device (operators, uploads, first step), several at a time, while the
// estimated device footprint of everything admitted fits the engine's
// budget.  A stepping thread keeps the admitted simulations in groups and
// advances each group with one launch per pass for all of its members:
// every block finds its problem in a per-pass tile table, each problem keeps
// its own FFT plans, tiling, medium, sources and sensors, and every per-step
// decision is taken on the device from the problem's own step counter, so
// members may be at different steps.  A group holds simulations with the
// same pass sequence (dimension, equation of state) and, where fixed-length
// kernels exist for the grid, the same size; simulations of other sizes
// share a group per pass sequence that runs the runtime-length kernels.
// Each group has at most one round of steps in flight, on whichever stream
// of a pool is free (least remaining work first), so groups overlap on the
// device and short simulations do not wait behind the rounds of long ones.
// Finished simulations are read
// back to host staging buffers and their device memory is released, which
// admits the next ones.
#include <algorithm>
#include <array>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>

#include "kw_sim.h"

namespace kw {
namespace {

using Clock = std::chrono::steady_clock;

// ---------------------------------------------------------------------------
// deep copy of a submitted problem
// ---------------------------------------------------------------------------
struct ProblemCopy {
  kw_problem pb{};
  std::vector<double> c0, rho0, rho0_sg[3], bona, alpha, alpha_filter, p0, tr, bc, dir_angle;
  std::vector<uint64_t> sens, tri, corners;
  struct Src {
    std::vector<uint64_t> index, sig_index;
    std::vector<double> input;
  } src[4];

  explicit ProblemCopy(const kw_problem& in) : pb(in) {
    const int64_t n = in.Nx * in.Ny * (in.dim == 3 ? in.Nz : 1);
    // grids are copied on several threads (a 256^3 problem carries about
    // 1 GB of them)
    auto grid = [&](const double* a, std::vector<double>& v) -> const double* {
      if (!a) return nullptr;
      v.resize((size_t)n);
      double* d = v.data();
      parallel_for(n, [&](int64_t i) { d[i] = a[i]; });
      return d;
    };
    pb.c0 = grid(in.c0, c0);
    pb.rho0 = grid(in.rho0, rho0);
    for (int d = 0; d < 3; ++d) pb.rho0_sg[d] = grid(in.rho0_sg[d], rho0_sg[d]);
    pb.BonA = grid(in.BonA, bona);
    pb.alpha_coeff = grid(in.alpha_coeff, alpha);
    pb.alpha_filter = grid(in.alpha_filter, alpha_filter);
    pb.p0 = grid(in.p0, p0);
    pb.sensor_mask_index = nullptr;
    if (in.sensor_mask_index && in.sensor_num_points > 0) {
      sens.assign(in.sensor_mask_index, in.sensor_mask_index + in.sensor_num_points);
      pb.sensor_mask_index = sens.data();
    }
    pb.sensor_mask_corners = nullptr;
    if (in.sensor_mask_corners && in.sensor_num_cuboids > 0) {
      corners.assign(in.sensor_mask_corners,
                     in.sensor_mask_corners + in.sensor_num_cuboids * 2 * in.dim);
      pb.sensor_mask_corners = corners.data();
    }
    pb.sensor_tri = nullptr;
    pb.sensor_bc = nullptr;
    if (in.sensor_tri && in.sensor_bc && in.sensor_num_points > 0) {
      const int64_t m = in.sensor_num_points * (in.dim + 1);
      tri.assign(in.sensor_tri, in.sensor_tri + m);
      bc.assign(in.sensor_bc, in.sensor_bc + m);
      pb.sensor_tri = tri.data();
      pb.sensor_bc = bc.data();
    }
    pb.directivity_angle = nullptr;
    if (in.directivity_angle && in.sensor_num_points > 0) {
      dir_angle.assign(in.directivity_angle, in.directivity_angle + in.sensor_num_points);
      pb.directivity_angle = dir_angle.data();
    }
    pb.time_reversal_boundary_data = nullptr;
    if (in.time_reversal_boundary_data && in.sensor_num_points > 0) {
      tr.assign(in.time_reversal_boundary_data,
                in.time_reversal_boundary_data + in.sensor_num_points * in.Nt);
      pb.time_reversal_boundary_data = tr.data();
    }
    copy_src(in.p_source, pb.p_source, src[0]);
    for (int d = 0; d < 3; ++d) copy_src(in.u_source[d], pb.u_source[d], src[1 + d]);
  }

  static void copy_src(const kw_time_source& a, kw_time_source& b, Src& s) {
    b = a;
    b.index = nullptr;
    b.sig_index = nullptr;
    b.input = nullptr;
    if (a.flag <= 0) return;
    if (a.index && a.num_points > 0) {
      s.index.assign(a.index, a.index + a.num_points);
      b.index = s.index.data();
    }
    if (a.sig_index && a.num_points > 0) {
      s.sig_index.assign(a.sig_index, a.sig_index + a.num_points);
      b.sig_index = s.sig_index.data();
    }
    if (a.input && a.num_series > 0) {
      s.input.assign(a.input, a.input + a.num_series * a.flag);
      b.input = s.input.data();
    }
  }
};

// the sizes kw_submit must trust to copy the arrays
void check_shape(const kw_problem* p) {
  if (!p) fail(KW_EINVAL, "problem is NULL");
  if (p->dim < 1 || p->dim > 3) fail(KW_EINVAL, "dim must be 1, 2 or 3");
  if (p->Nx < 2 || (p->dim == 1 && (p->Ny != 1 || p->Nz != 1)) ||
      (p->dim >= 2 && p->Ny < 2) || (p->dim == 3 && p->Nz < 2) || (p->dim == 2 && p->Nz != 1))
    fail(KW_EINVAL, "invalid grid size");
  if ((double)p->Nx * (double)p->Ny * (double)(p->dim == 3 ? p->Nz : 1) > 4.0e9)
    fail(KW_EUNSUP, "grid too large");
  if (p->Nt < 1) fail(KW_EINVAL, "Nt must be positive");
  if (p->sensor_num_points < 0) fail(KW_EINVAL, "negative number of sensor points");
  if (p->sensor_num_cuboids < 0) fail(KW_EINVAL, "negative number of sensor cuboids");
  const kw_time_source* s[4] = {&p->p_source, &p->u_source[0], &p->u_source[1], &p->u_source[2]};
  for (auto* t : s)
    if (t->flag > 0 && (t->num_points < 0 || t->num_series < 0))
      fail(KW_EINVAL, "invalid source sizes");
}

// device bytes a problem will hold (the allocations of create_sim, with
// margin); used to admit simulations before they are created
int64_t estimate_bytes(const kw_problem& p) {
  // a 1-D problem runs as a one-row 2-D grid
  const int dim = std::max(p.dim, 2);
  const Geometry G = make_geometry(dim, p.Nx, p.Ny, p.Nz);
  const double nc = (double)G.ncell, ns = (double)G.nspec;
  double b = 4.0 * nc * (2 + 2 * dim) + 8.0 * ns * 3 + 4.0 * ns * 3;
  auto grid = [&](const void* a) { return a ? 4.0 * nc : 0.0; };
  b += grid(p.c0) + grid(p.rho0) + grid(p.BonA) + grid(p.p0);
  if (p.rho0 || p.rho0_sg[0]) b += 4.0 * nc * dim;
  if (p.absorbing_flag && (p.c0 || p.alpha_coeff)) b += 8.0 * nc;
  const kw_time_source* srcs[4] = {&p.p_source, &p.u_source[0], &p.u_source[1], &p.u_source[2]};
  for (auto* s : srcs) {
    if (s->flag <= 0) continue;
    b += 8.0 * nc + 4.0 * (double)s->num_series * (double)std::min<int64_t>(s->flag, p.Nt) +
         8.0 * (double)s->num_points;
    if (s->mode == KW_SRC_ADDITIVE) b += 4.0 * ns;
  }
  const double nsens = (double)p.sensor_num_points;
  if (p.time_reversal_boundary_data) {
    b += 4.0 * nc + 4.0 * nsens * (double)p.Nt;
  } else if (nsens > 0) {
    const int64_t rs = p.record_start_index > 0 ? p.record_start_index : 1;
    const double nrec = (double)std::max<int64_t>(0, p.Nt - rs + 1);
    const uint32_t r = p.record_flags;
    if (r & KW_REC_P) b += 4.0 * nsens * nrec;
    if (r & KW_REC_U) b += 4.0 * nsens * nrec * dim;
    for (uint32_t f : {KW_REC_P_MAX, KW_REC_P_MIN, KW_REC_P_RMS})
      if (r & f) b += 8.0 * nsens;
    for (uint32_t f : {KW_REC_U_MAX, KW_REC_U_MIN, KW_REC_U_RMS})
      if (r & f) b += 8.0 * nsens * dim;
    for (uint32_t f : {KW_REC_P_MAX_ALL, KW_REC_P_MIN_ALL})
      if (r & f) b += 8.0 * nc;
    for (uint32_t f : {KW_REC_U_MAX_ALL, KW_REC_U_MIN_ALL})
      if (r & f) b += 8.0 * nc * dim;
    b += 4.0 * nsens;
  }
  return (int64_t)(b * 1.1) + ((int64_t)4 << 20);
}

// ---------------------------------------------------------------------------
// host staging of a finished simulation's outputs
// ---------------------------------------------------------------------------
struct OutputStage {
  std::vector<float> p, p_final, u[3], u_final[3], u_ns[3], u_sp[3], u_ss[3];
  std::vector<double> p_max, p_min, p_rms, p_max_all, p_min_all;
  std::vector<double> u_max[3], u_min[3], u_rms[3], u_max_all[3], u_min_all[3];

  void read(Sim& S) {
    const int64_t ns = S.nsens, nc = S.G.ncell, nr = S.n_rec;
    const uint32_t r = S.record;
    kw_outputs o{};
    auto f32 = [](std::vector<float>& v, int64_t n, bool on) -> float* {
      if (!on) return nullptr;
      v.assign(n, 0.f);
      return v.data();
    };
    auto f64 = [](std::vector<double>& v, int64_t n, bool on) -> double* {
      if (!on) return nullptr;
      v.assign(n, 0.0);
      return v.data();
    };
    o.p = f32(p, ns * nr, (r & KW_REC_P) && ns > 0);
    o.p_max = f64(p_max, ns, (r & KW_REC_P_MAX) && ns > 0);
    o.p_min = f64(p_min, ns, (r & KW_REC_P_MIN) && ns > 0);
    o.p_rms = f64(p_rms, ns, (r & KW_REC_P_RMS) && ns > 0);
    o.p_max_all = f64(p_max_all, nc, r & KW_REC_P_MAX_ALL);
    o.p_min_all = f64(p_min_all, nc, r & KW_REC_P_MIN_ALL);
    o.p_final = f32(p_final, nc, r & KW_REC_P_FINAL);
    for (int a = 0; a < S.G.dim; ++a) {
      o.u[a] = f32(u[a], ns * nr, (r & KW_REC_U) && ns > 0);
      o.u_max[a] = f64(u_max[a], ns, (r & KW_REC_U_MAX) && ns > 0);
      o.u_min[a] = f64(u_min[a], ns, (r & KW_REC_U_MIN) && ns > 0);
      o.u_rms[a] = f64(u_rms[a], ns, (r & KW_REC_U_RMS) && ns > 0);
      o.u_max_all[a] = f64(u_max_all[a], nc, r & KW_REC_U_MAX_ALL);
      o.u_min_all[a] = f64(u_min_all[a], nc, r & KW_REC_U_MIN_ALL);
      o.u_final[a] = f32(u_final[a], nc, r & KW_REC_U_FINAL);
      o.u_non_staggered[a] = f32(u_ns[a], ns * nr, (r & KW_REC_U_NON_STAGGERED) && ns > 0);
      o.u_split_p[a] = f32(u_sp[a], ns * nr, (r & KW_REC_U_SPLIT_FIELD) && ns > 0);
      o.u_split_s[a] = f32(u_ss[a], ns * nr, (r & KW_REC_U_SPLIT_FIELD) && ns > 0);
    }
    read_outputs(S, &o);
  }

  // copy into the caller's buffers (NULL pointers and outputs the problem
  // did not record are skipped, as kw_run does)
  void copy_to(kw_outputs* o) const {
    if (!o) return;
    auto c = [](auto* dst, const auto& v) {
      if (dst && !v.empty()) std::memcpy(dst, v.data(), v.size() * sizeof(v[0]));
    };
    c(o->p, p);
    c(o->p_max, p_max);
    c(o->p_min, p_min);
    c(o->p_rms, p_rms);
    c(o->p_max_all, p_max_all);
    c(o->p_min_all, p_min_all);
    c(o->p_final, p_final);
    for (int a = 0; a < 3; ++a) {
      c(o->u[a], u[a]);
      c(o->u_max[a], u_max[a]);
      c(o->u_min[a], u_min[a]);
      c(o->u_rms[a], u_rms[a]);
      c(o->u_max_all[a], u_max_all[a]);
      c(o->u_min_all[a], u_min_all[a]);
      c(o->u_final[a], u_final[a]);
      c(o->u_non_staggered[a], u_ns[a]);
      c(o->u_split_p[a], u_sp[a]);
      c(o->u_split_s[a], u_ss[a]);
    }
  }
};

// ---------------------------------------------------------------------------
// groups: one launch per pass for all members
// ---------------------------------------------------------------------------
enum GeoKind { GX = 0, GY, GL, GCELL, GSENS, NGEO };

struct Ticket;

struct Group {
  int dim = 0;
  bool absorbing = false;
  bool axisym = false;  // axisymmetric members: stepped one by one
  void* stream = nullptr;
  void* ev = nullptr;       // recorded after the last round issued
  bool in_flight = false;   // that round has not finished
  void* graph = nullptr;    // captured round of graph_k steps of this membership
  int64_t graph_k = 0, graph_launches = 0;
  std::vector<Ticket*> members, joining;
  std::vector<Sim*> sims;
  bool dirty = true;
  bool changed = true;  // membership differs from the one the tables describe
  int slot = -1;        // stream of the pool its round in flight uses
  ProbDev* probs = nullptr;
  int32_t* tiles[NGEO] = {};
  LaunchGeom geo[NGEO];
  int fixed[3] = {0, 0, 0};  // x, y, last-axis pass lengths (0: runtime plans)
  bool any_points = false, any_grid = false;
  std::vector<void*> live, retired;  // device tables of this / earlier memberships
  int64_t launches = 0;

  Group() = default;
  Group(const Group&) = delete;
  Group& operator=(const Group&) = delete;
  ~Group() {
    dev_graph_destroy(graph);
    for (void* p : live) dev_free(p);
    for (void* p : retired) dev_free(p);
    if (ev) dev_event_destroy(ev);
  }

  // a round can be captured once and replayed when nothing in it depends on
  // host state that changes from step to step (the per-step source grids of
  // several-series additive sources are chosen on the host)
  bool graphable() const {
    // an axisymmetric member's step decides on the host what it records and
    // whether it stores p for p_final, which a replayed round would not redo
    if (axisym) return false;
    for (Sim* s : sims)
      if (s->p_grid || s->u_grid[0] || s->u_grid[1] || s->u_grid[2]) return false;
    return true;
  }

  // tables of earlier memberships may be read by launches still queued;
  // free them only after the stream has been synchronised
  void free_retired() {
    for (void* p : retired) dev_free(p);
    retired.clear();
  }

  static const LaunchGeom& geom_of(const Sim& S, int k) {
    switch (k) {
      case GX: return S.lx;
      case GY: return S.ly;
      case GL: return S.ll;
      case GCELL: return S.lcell;
      default: return S.lsens;
    }
  }

  void rebuild() {
    dev_graph_destroy(graph);  // called between rounds: the stream is idle
    graph = nullptr;
    graph_k = 0;
    for (void* p : live) retired.push_back(p);
    live.clear();
    const int n = (int)sims.size();
    if (n == 0) {
      dirty = false;
      return;
    }
    std::vector<ProbDev> hp(n);
    for (int i = 0; i < n; ++i) hp[i] = sims[i]->P;
    probs = (ProbDev*)dev_alloc(sizeof(ProbDev) * n);
    live.push_back(probs);
    dev_h2d(probs, hp.data(), sizeof(ProbDev) * n, stream);
    for (int k = 0; k < NGEO; ++k) {
      if (k == GY && dim != 3) {
        tiles[k] = nullptr;
        geo[k] = LaunchGeom{};
        continue;
      }
      std::vector<int32_t> start(n);
      int64_t nb = 0;
      int T = 0;
      for (int i = 0; i < n; ++i) {
        const LaunchGeom& g = geom_of(*sims[i], k);
        start[i] = (int32_t)nb;
        nb += g.nb;
        T = std::max(T, g.T);
      }
      if (nb >= INT32_MAX) fail(KW_EUNSUP, "too many blocks in one group launch");
      // a block has T threads for every member: an x pass member gives each
      // of its T / Gx pencil slots its own LDS pencil; the strided and
      // last-axis passes keep their own pencil count
      size_t lds = 0;
      for (int i = 0; i < n; ++i) {
        const Sim& S = *sims[i];
        size_t l = geom_of(S, k).lds;
        if (k == GX)
          l = sizeof(cf) * ((size_t)kNPX * (size_t)x_lds_stride(S.P.fx.L) *
                                (size_t)((T + S.P.Gx - 1) / S.P.Gx) +
                            (size_t)S.P.fx.L);
        lds = std::max(lds, l);
      }
      if (lds > dev_max_lds()) fail(KW_EUNSUP, "group launch needs too much LDS");
      geo[k] = LaunchGeom{(int)nb, T, lds};
      tiles[k] = (int32_t*)dev_alloc(sizeof(int32_t) * n);
      live.push_back(tiles[k]);
      dev_h2d(tiles[k], start.data(), sizeof(int32_t) * n, stream);
    }
    // a fixed-length variant serves the launch only if every member has
    // that length on the pass's axis
    auto common = [&](auto len, PassKind kind) -> int {
      const int N = len(*sims[0]);
      for (Sim* s : sims)
        if (len(*s) != N) return 0;
      return std::getenv("KW_NO_FIXED") ? 0 : fixed_variant(N, kind);
    };
    fixed[0] = common([](const Sim& s) { return s.G.Nx; }, PK_X_GRAD);
    fixed[1] = dim == 3 ? common([](const Sim& s) { return s.G.Ny; }, PK_YFWD) : 0;
    fixed[2] = common([&](const Sim& s) { return dim == 3 ? s.G.Nz : s.G.Ny; }, PK_SPEC);
    any_points = any_grid = false;
    for (Sim* s : sims) {
      any_points = any_points || s->P.rec_points != 0;
      any_grid = any_grid || s->P.rec_grid != 0;
    }
    dirty = false;
  }

  Batch batch(int k) const { return Batch{probs, tiles[k], (int32_t)sims.size()}; }

  void pass(PassKind kind, const PassArgs& a0) {
    const int k = (kind == PK_YFWD || kind == PK_YINV) ? GY : (kind == PK_SPEC ? GL : GX);
    const int fx = fixed[k == GX ? 0 : (k == GY ? 1 : 2)];
    PassArgs a = a0;
    int nb = geo[k].nb;
    if (k == GY && a.nbuf > 1 && ysplit_on()) {
      a.ysplit = nb;
      nb *= a.nbuf;
    }
    launch_pass(kind, fx, dim, nb, geo[k].T, launch_lds(kind, fx, a, geo[k].lds), batch(k), a,
                stream);
    ++launches;
  }

  void ew(const EwArgs& a, int k) {
    launch_ew(geo[k].nb, batch(k), a, stream);
    ++launches;
  }
};

// one time step (t >= 2) of every member: the sequence of step_once for a
// single simulation, with each pass launched once for the group
void group_step(Group& g) {
  if (g.axisym) {
    // the u_r / r term, the mirror kernels and the last radial column of
    // 'WSWS-FFT' are single-simulation launches: members step one by one
    // on the group's stream
    for (Sim* s : g.sims) {
      step_once(*s);
      g.launches += 16;
    }
    return;
  }
  for (Sim* s : g.sims) {
    const int t = (int)(s->steps_done + 1);
    if (s->p_grid) src_grid(*s, 0, t);
    for (int a = 0; a < g.dim; ++a)
      if (s->u_grid[a]) src_grid(*s, a + 1, t);
  }
  const int dim = g.dim;
  // the first pass of the step advances every member's step number
  PassArgs y = bufs(1, 0);
  y.advance = 1;
  if (dim == 3) g.pass(PK_YFWD, y);
  PassArgs gr{};
  gr.mode = SM_GRAD;
  gr.advance = dim == 3 ? 0 : 1;
  g.pass(PK_SPEC, gr);
  if (dim == 3) g.pass(PK_YINV, bufs(3));
  g.pass(PK_X_GRAD, PassArgs{});
  if (dim == 3) g.pass(PK_YFWD, bufs(3));
  PassArgs dv{};
  dv.mode = SM_DIV;
  g.pass(PK_SPEC, dv);
  if (dim == 3) g.pass(PK_YINV, bufs(3));
  g.pass(PK_X_DIV, PassArgs{});
  if (g.absorbing) {
    if (dim == 3) g.pass(PK_YFWD, bufs(2));
    PassArgs ab{};
    ab.mode = SM_ABS;
    g.pass(PK_SPEC, ab);
    if (dim == 3) g.pass(PK_YINV, bufs(2));
    g.pass(PK_X_EOS, PassArgs{});
  }
  if (g.any_points) {
    EwArgs e{};
    e.kind = EW_RECORD_POINTS;
    e.rec = kRecAuto;
    g.ew(e, GSENS);
  }
  if (g.any_grid) {
    EwArgs e{};
    e.kind = EW_RECORD_GRID;
    e.rec = kRecAuto;
    g.ew(e, GCELL);
  }
  // non-staggered velocity: per member, on the group's stream (its gather
  // decides on the device whether the step is recorded)
  for (Sim* s : g.sims)
    if (s->record & (KW_REC_U_NON_STAGGERED | KW_REC_U_SPLIT_FIELD)) {
      record_velocity(*s);
      g.launches += (g.dim == 3 ? 6 : 4) * g.dim;
      if (s->record & KW_REC_U_SPLIT_FIELD) g.launches += g.dim == 3 ? 9 : 5;
    }
  for (Sim* s : g.sims)
    if (!s->dir_tab.empty()) {
      record_directivity(*s);
      g.launches += 4 * (int64_t)s->dir_tab.size();
    }
  for (Sim* s : g.sims) s->steps_done++;
}

// ---------------------------------------------------------------------------
// tickets and the engine
// ---------------------------------------------------------------------------
struct Ticket {
  int64_t id = 0;
  std::unique_ptr<ProblemCopy> prob;
  int32_t status = KW_TICKET_QUEUED;
  int32_t error = 0;
  std::string msg;
  double submit_s = 0, start_s = -1, done_s = -1;
  int64_t est_bytes = 0, bytes = 0, steps_total = 0, steps_done = 0, cells = 0;
  bool cancel = false;
  int flag = -1;             // slot of its failure flag while on the device
  std::unique_ptr<Sim> sim;  // owned by the stepping thread while running
  OutputStage out;
};

bool terminal(int32_t s) {
  return s == KW_TICKET_DONE || s == KW_TICKET_FAILED || s == KW_TICKET_CANCELLED;
}

}  // namespace

struct Engine {
  std::mutex mu;
  std::condition_variable cv_admit, cv_step, cv_done;
  std::deque<Ticket*> pending;  // submitted, not yet admitted
  std::vector<Ticket*> ready;   // on the device after their first step
  std::map<int64_t, std::unique_ptr<Ticket>> tickets;
  int64_t next_id = 1;
  int64_t budget = 0, held = 0;  // device bytes: budget, held (estimates until created)
  int32_t steps_per_round = 32;
  bool use_graphs = true;
  bool sjf_only = false;  // one lane: least remaining work first
  // cells per launch above which a group's launches fill the device
  static constexpr double kBigCells = double(1 << 22);
  bool stop = false;
  std::vector<void*> streams;  // compute streams, one per group (round robin)
  void* admit_stream = nullptr;
  Clock::time_point t0 = Clock::now();
  // statistics (under mu)
  int64_t n_submitted = 0, n_completed = 0, n_failed = 0, n_cancelled = 0, n_running = 0;
  double cell_steps = 0, busy_s = 0;
  int64_t launches = 0;
  std::vector<std::thread> admit_thr;
  std::thread step_thr;
  // per-simulation failure flags in pinned host memory: the kernels that
  // form the pressure write the first step with a non-finite value into the
  // simulation's slot, and the stepping thread reads it after each round
  // without a device copy
  static constexpr int kFlagSlots = 1 << 16;
  int32_t* flags = nullptr;
  std::vector<int> free_flags;  // under mu

  double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }

  int32_t flag_of(const Ticket* t) const {
    return t->flag < 0 ? 0 : __atomic_load_n(&flags[t->flag], __ATOMIC_ACQUIRE);
  }
  void release_flag(Ticket* t) {
    // under mu
    if (t->flag >= 0) free_flags.push_back(t->flag);
    t->flag = -1;
  }

  void finish(Ticket* t, int32_t status, int32_t err, const std::string& m) {
    // under mu
    t->status = status;
    t->error = err;
    t->msg = m;
    t->done_s = now();
    if (status == KW_TICKET_DONE) ++n_completed;
    if (status == KW_TICKET_FAILED) ++n_failed;
    if (status == KW_TICKET_CANCELLED) ++n_cancelled;
    cv_done.notify_all();
  }

  // -- admission ------------------------------------------------------------
  void admit_loop() {
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
      cv_admit.wait(lk, [&]() {
        if (stop) return true;
        if (pending.empty()) return false;
        const Ticket* t = pending.front();
        // admit while the budget holds; always admit when nothing is held
        return t->cancel || held == 0 || held + t->est_bytes <= budget;
      });
      if (stop) return;
      Ticket* t = pending.front();
      pending.pop_front();
      if (t->cancel) {
        finish(t, KW_TICKET_CANCELLED, KW_ESTATE, "cancelled");
        continue;
      }
      if (free_flags.empty()) {
        finish(t, KW_TICKET_FAILED, KW_ENOMEM, "too many simulations on the device at once");
        continue;
      }
      t->flag = free_flags.back();
      free_flags.pop_back();
      held += t->est_bytes;
      t->status = KW_TICKET_RUNNING;
      t->start_s = now();
      ++n_running;
      lk.unlock();
      std::unique_ptr<Sim> sim;
      int code = KW_OK;
      std::string msg;
      try {
        {
          // tiles come from the autotuner's cache (or the built-in rule):
          // a measurement here would hold up every admission behind it
          TuneCacheOnly cache_only;
          sim = create_sim(&t->prob->pb, admit_stream);
        }
        // the failure flag lives in the engine's pinned slot
        __atomic_store_n(&flags[t->flag], 0, __ATOMIC_RELEASE);
        sim->P.bad = flags + t->flag;
        dev_h2d(sim->dP, &sim->P, sizeof(ProbDev), admit_stream);
        if (sim->Nt_eff > 0) step_once(*sim);
        dev_sync(admit_stream);
      } catch (const DeviceError& e) {
        code = e.code;
        msg = e.what();
      } catch (const std::bad_alloc&) {
        code = KW_ENOMEM;
        msg = "host allocation failed";
      } catch (const std::exception& e) {
        code = KW_EDEVICE;
        msg = e.what();
      }
      lk.lock();
      // the estimate is replaced by what the simulation holds: another
      // admission worker may now fit
      cv_admit.notify_all();
      if (code != KW_OK) {
        sim.reset();
        held -= t->est_bytes;
        --n_running;
        release_flag(t);
        finish(t, KW_TICKET_FAILED, code, msg);
        continue;
      }
      held += sim->bytes - t->est_bytes;
      t->bytes = sim->bytes;
      t->steps_total = sim->Nt_eff;
      t->steps_done = sim->steps_done;
      t->cells = sim->G.ncell;
      t->prob.reset();  // the device holds everything now
      t->sim = std::move(sim);
      ready.push_back(t);
      cv_step.notify_one();
    }
  }

  // -- stepping -------------------------------------------------------------
  void retire(Ticket* t, std::unique_lock<std::mutex>& lk, bool cancelled) {
    // called without mu held for the device work
    std::string msg;
    int code = KW_OK;
    if (!cancelled) {
      try {
        t->out.read(*t->sim);
      } catch (const DeviceError& e) {
        code = e.code;
        msg = e.what();
      }
    } else {
      try {
        dev_sync(t->sim->stream);
      } catch (...) {
      }
    }
    const int64_t b = t->bytes;
    t->sim.reset();
    lk.lock();
    held -= b;
    --n_running;
    release_flag(t);
    if (cancelled)
      finish(t, KW_TICKET_CANCELLED, KW_ESTATE, "cancelled");
    else if (code != KW_OK)
      finish(t, KW_TICKET_FAILED, code, msg);
    else
      finish(t, KW_TICKET_DONE, KW_OK, "");
    cv_admit.notify_one();
    lk.unlock();
  }

  // a simulation whose pressure became non-finite at step `bad` fails alone
  // (called without mu held; its group's round has ended)
  void drop_diverged(Ticket* t, int32_t bad, std::unique_lock<std::mutex>& lk) {
    const int64_t b = t->bytes;
    t->sim.reset();
    lk.lock();
    held -= b;
    --n_running;
    release_flag(t);
    finish(t, KW_TICKET_FAILED, KW_ENUMERIC,
           "the pressure became non-finite at time step " + std::to_string(bad) +
               " (the simulation is unstable: check the time step against the stability limit)");
    cv_admit.notify_one();
    lk.unlock();
  }

  // Group of a simulation: same pass sequence and, when fixed-length
  // variants exist for every axis of its grid, the same grid size (the group
  // then runs those variants); other sizes share one group per pass sequence
  // and run the runtime-length variants.
  static std::array<int, 6> group_key(const Sim& s) {
    const int dim = s.G.dim;
    const int absorbing = s.P.eos == EOS_ABSORBING ? 1 : 0;
    const int Nl = dim == 3 ? s.G.Nz : s.G.Ny;
    const bool fixed = !std::getenv("KW_NO_FIXED") && fixed_variant(s.G.Nx, PK_X_GRAD) != 0 &&
                       (dim == 2 || fixed_variant(s.G.Ny, PK_YFWD) != 0) &&
                       fixed_variant(Nl, PK_SPEC) != 0;
    // axisymmetric simulations (1 'WSWA-FFT', 3 'WSWS-FFT') group with
    // their own kind and size only
    const int as = s.axisym ? (s.as_ws ? 3 : 1) : 0;
    if (fixed || as) return {dim, absorbing, s.G.Nx, s.G.Ny, s.G.Nz, as};
    return {dim, absorbing, 0, 0, 0, 0};
  }

  // A group has at most one round (steps_per_round steps of all members) in
  // flight, on whichever stream of the pool is free when the round is
  // issued; rounds of different groups overlap on the device.  Members join
  // and leave a group only between its rounds.  Free streams go to groups in
  // two lanes (large launches and small ones, each by least remaining work),
  // so short simulations are not queued behind the rounds of long ones and
  // large groups keep the device busy while small ones wait on latency.
  void step_loop() {
    std::map<std::array<int, 6>, std::unique_ptr<Group>> groups;
    std::vector<char> stream_busy(streams.size(), 0);
    auto last = Clock::now();
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
      bool any = false;
      for (auto& kv : groups) any = any || !kv.second->members.empty() || !kv.second->joining.empty();
      if (!any) cv_step.wait(lk, [&]() { return stop || !ready.empty(); });
      if (stop) break;
      std::vector<Ticket*> joining(ready.begin(), ready.end());
      ready.clear();
      lk.unlock();
      for (Ticket* t : joining) {
        if (const int32_t bad = flag_of(t)) {
          drop_diverged(t, bad, lk);
          continue;
        }
        if (t->sim->steps_done >= t->sim->Nt_eff) {
          retire(t, lk, false);  // finished in its first step
          continue;
        }
        const auto key = group_key(*t->sim);
        auto& g = groups[key];
        if (!g) {
          g = std::make_unique<Group>();
          g->dim = key[0];
          g->absorbing = key[1] != 0;
          g->axisym = key[5] != 0;
          g->ev = dev_event_create();
        }
        g->joining.push_back(t);
      }
      bool issued = false, busy = false;
      double cells = 0;
      int64_t nl = 0;
      // rounds that have ended free their streams; between rounds, members
      // that finished or were cancelled leave and new ones join
      std::vector<std::pair<double, Group*>> idle;
      for (auto& kv : groups) {
        Group& g = *kv.second;
        if (g.in_flight) {
          if (!dev_event_done(g.ev)) {
            busy = true;
            continue;
          }
          g.in_flight = false;
          stream_busy[g.slot] = 0;
        }
        std::vector<Ticket*> keep;
        for (Ticket* t : g.members) {
          lk.lock();
          const bool cancelled = t->cancel;
          lk.unlock();
          const int32_t bad = cancelled ? 0 : flag_of(t);
          if (bad)
            drop_diverged(t, bad, lk);
          else if (cancelled || t->sim->steps_done >= t->sim->Nt_eff)
            retire(t, lk, cancelled);
          else
            keep.push_back(t);
        }
        g.changed = g.changed || keep.size() != g.members.size() || !g.joining.empty();
        for (Ticket* t : g.joining) keep.push_back(t);
        g.joining.clear();
        g.members.swap(keep);
        if (g.members.empty()) continue;
        double work = 0;
        for (Ticket* t : g.members)
          work += (double)t->sim->G.ncell * (double)(t->sim->Nt_eff - t->sim->steps_done);
        idle.push_back({work, &g});
      }
      std::sort(idle.begin(), idle.end(),
                [](const std::pair<double, Group*>& a, const std::pair<double, Group*>& b) {
                  return a.first < b.first;
                });
      // two lanes: the first half of the streams serves groups whose
      // launches fill the device, the other half the latency-bound small
      // groups, each by least remaining work first; a lane with nothing of
      // its own kind takes the other kind, so small rounds run beside large
      // ones instead of queueing in front of them
      std::vector<Group*> big, small;
      for (auto& wg : idle) {
        double cells = 0;
        for (Ticket* t : wg.second->members) cells += (double)t->sim->G.ncell;
        (sjf_only || cells < kBigCells ? small : big).push_back(wg.second);
      }
      size_t ib = 0, is = 0;
      const size_t nbig_lanes = sjf_only ? 0 : (streams.size() + 1) / 2;
      for (;;) {
        if (ib == big.size() && is == small.size()) break;
        int slot = -1;
        for (size_t k = 0; k < streams.size() && slot < 0; ++k)
          if (!stream_busy[k]) slot = (int)k;
        if (slot < 0) {
          busy = true;  // every stream has a round in flight
          break;
        }
        const bool big_lane = (size_t)slot < nbig_lanes;
        Group* pick;
        if (big_lane ? ib < big.size() : is == small.size())
          pick = big[ib++];
        else
          pick = small[is++];
        Group& g = *pick;
        int code = KW_OK;
        std::string err;
        try {
          g.slot = slot;
          g.stream = streams[slot];
          for (Ticket* t : g.members) t->sim->stream = g.stream;
          if (g.changed) {
            g.sims.clear();
            for (Ticket* t : g.members) g.sims.push_back(t->sim.get());
            g.free_retired();  // the group's last round has ended
            g.rebuild();
            g.changed = false;
          }
          int64_t k = steps_per_round;
          for (Sim* s : g.sims) k = std::min(k, s->Nt_eff - s->steps_done);
          const int64_t l0 = g.launches;
          if (use_graphs && k == steps_per_round && g.graphable()) {
            // full rounds of a membership are captured once and replayed
            // (on any stream): every kernel takes its step from the device,
            // so the launches of one round are the same for every round
            if (!g.graph && dev_capture_begin(g.stream)) {
              try {
                for (int64_t i = 0; i < k; ++i) group_step(g);
              } catch (...) {
                dev_capture_abort(g.stream);
                throw;
              }
              g.graph = dev_capture_end(g.stream);
              g.graph_k = k;
              g.graph_launches = g.launches - l0;
              dev_graph_launch(g.graph, g.stream);
            } else if (g.graph) {
              dev_graph_launch(g.graph, g.stream);
              for (Sim* s : g.sims) s->steps_done += g.graph_k;
              g.launches += g.graph_launches;
            } else {
              for (int64_t i = 0; i < k; ++i) group_step(g);
            }
          } else {
            for (int64_t i = 0; i < k; ++i) group_step(g);
          }
          dev_event_record(g.ev, g.stream);
          g.in_flight = true;
          stream_busy[slot] = 1;
          issued = busy = true;
          nl += g.launches - l0;
          for (Sim* s : g.sims) cells += (double)s->G.ncell * (double)k;
        } catch (const DeviceError& e) {
          code = e.code;
          err = e.what();
        }
        if (code != KW_OK) {
          // a device error ends every member of the group
          try {
            dev_sync(g.stream);
          } catch (...) {
          }
          g.in_flight = false;
          stream_busy[slot] = 0;
          std::vector<Ticket*> all = g.members;
          g.members.clear();
          g.sims.clear();
          g.changed = true;
          for (Ticket* t : all) {
            const int64_t b = t->bytes;
            t->sim.reset();
            lk.lock();
            held -= b;
            --n_running;
            release_flag(t);
            finish(t, KW_TICKET_FAILED, code, err);
            lk.unlock();
          }
        }
      }
      // groups left without members and without a round in flight go
      for (auto it = groups.begin(); it != groups.end();) {
        Group& g = *it->second;
        if (g.members.empty() && g.joining.empty() && !g.in_flight)
          it = groups.erase(it);
        else
          ++it;
      }
      const auto nowt = Clock::now();
      lk.lock();
      for (auto& kv : groups)
        for (Ticket* t : kv.second->members) t->steps_done = t->sim->steps_done;
      cell_steps += cells;
      launches += nl;
      if (busy) busy_s += std::chrono::duration<double>(nowt - last).count();
      last = nowt;
      if (!issued && busy) {
        // every group has a round in flight: wait a little for one to end
        // (or for new simulations)
        cv_step.wait_for(lk, std::chrono::microseconds(100),
                         [&]() { return stop || !ready.empty(); });
      }
    }
    // shutting down: release everything still on the device
    lk.unlock();
    for (void* st : streams) {
      try {
        dev_sync(st);
      } catch (...) {
      }
    }
    std::vector<Ticket*> all;
    for (auto& kv : groups) {
      all.insert(all.end(), kv.second->members.begin(), kv.second->members.end());
      all.insert(all.end(), kv.second->joining.begin(), kv.second->joining.end());
    }
    groups.clear();
    lk.lock();
    all.insert(all.end(), ready.begin(), ready.end());
    ready.clear();
    for (Ticket* t : all) {
      t->sim.reset();
      held -= t->bytes;
      --n_running;
      release_flag(t);
      finish(t, KW_TICKET_CANCELLED, KW_ESTATE, "engine destroyed");
    }
  }
};

}  // namespace kw

struct kw_engine {
  kw::Engine e;
};

extern "C" {

int kw_engine_create(const kw_engine_config* cfg, kw_engine** engine, char* err, size_t errlen) {
  return kw_guarded(err, errlen, [&]() {
    if (!engine) kw::fail(KW_EINVAL, "NULL engine pointer");
    *engine = nullptr;
    auto h = std::make_unique<kw_engine>();
    kw::Engine& E = h->e;
    size_t fr = 0, tot = 0;
    kw::dev_mem_info(&fr, &tot);
    E.budget = (cfg && cfg->device_bytes > 0) ? cfg->device_bytes : (int64_t)((double)fr * 0.9);
    E.steps_per_round = (cfg && cfg->steps_per_round > 0) ? cfg->steps_per_round : 32;
    E.use_graphs = std::getenv("KW_NO_GRAPHS") == nullptr;
    E.sjf_only = std::getenv("KW_ENGINE_SJF") != nullptr;
    E.flags = (int32_t*)kw::dev_host_alloc(sizeof(int32_t) * kw::Engine::kFlagSlots);
    std::memset(E.flags, 0, sizeof(int32_t) * kw::Engine::kFlagSlots);
    E.free_flags.reserve(kw::Engine::kFlagSlots);
    for (int i = kw::Engine::kFlagSlots - 1; i >= 0; --i) E.free_flags.push_back(i);
    // as many compute streams as the device has hardware queues for this
    // process by default (4), so that groups overlap without oversubscribing
    for (int i = 0; i < 4; ++i) E.streams.push_back(kw::dev_stream_create());
    // admission workers share one stream: their device work (uploads, the
    // first step) is short next to the host-side operator construction they
    // overlap
    E.admit_stream = kw::dev_stream_create();
    const int na = (cfg && cfg->admit_threads > 0) ? std::min(cfg->admit_threads, 16) : 3;
    for (int i = 0; i < na; ++i) E.admit_thr.emplace_back([&E]() { E.admit_loop(); });
    E.step_thr = std::thread([&E]() { E.step_loop(); });
    *engine = h.release();
  });
}

int kw_submit(kw_engine* engine, const kw_problem* prob, int64_t* ticket, char* err,
              size_t errlen) {
  return kw_guarded(err, errlen, [&]() {
    if (!engine || !ticket) kw::fail(KW_EINVAL, "NULL argument");
    kw::check_shape(prob);
    if (prob->axisymmetric == 2)
      kw::fail(KW_EINVAL, "axisymmetric = 2 (the radial grid of the file contract) is run by "
                          "kw_run_file");
    auto t = std::make_unique<kw::Ticket>();
    t->prob = std::make_unique<kw::ProblemCopy>(*prob);
    t->est_bytes = kw::estimate_bytes(*prob);
    kw::Engine& E = engine->e;
    std::lock_guard<std::mutex> lk(E.mu);
    if (E.stop) kw::fail(KW_ESTATE, "engine is shutting down");
    t->id = E.next_id++;
    t->submit_s = E.now();
    t->cells = prob->Nx * prob->Ny * (prob->dim == 3 ? prob->Nz : 1);
    *ticket = t->id;
    E.pending.push_back(t.get());
    E.tickets[t->id] = std::move(t);
    ++E.n_submitted;
    E.cv_admit.notify_one();
  });
}

int kw_poll(kw_engine* engine, int64_t ticket, kw_ticket_info* info, char* err, size_t errlen) {
  return kw_guarded(err, errlen, [&]() {
    if (!engine || !info) kw::fail(KW_EINVAL, "NULL argument");
    kw::Engine& E = engine->e;
    std::lock_guard<std::mutex> lk(E.mu);
    auto it = E.tickets.find(ticket);
    if (it == E.tickets.end()) kw::fail(KW_EINVAL, "unknown ticket");
    const kw::Ticket& t = *it->second;
    info->status = t.status;
    info->error = t.error;
    info->submit_s = t.submit_s;
    info->start_s = t.start_s;
    info->done_s = t.done_s;
    info->steps_done = t.steps_done;
    info->steps_total = t.steps_total;
    info->grid_cells = t.cells;
    info->device_bytes = t.bytes;
  });
}

int kw_wait(kw_engine* engine, int64_t ticket, kw_outputs* out, char* err, size_t errlen) {
  if (!engine) {
    kw_set_err(err, errlen, "NULL engine");
    return KW_EINVAL;
  }
  kw::Engine& E = engine->e;
  std::unique_ptr<kw::Ticket> t;
  {
    std::unique_lock<std::mutex> lk(E.mu);
    auto it = E.tickets.find(ticket);
    if (it == E.tickets.end()) {
      kw_set_err(err, errlen, "unknown ticket");
      return KW_EINVAL;
    }
    kw::Ticket* tp = it->second.get();
    E.cv_done.wait(lk, [&]() { return kw::terminal(tp->status); });
    t = std::move(it->second);
    E.tickets.erase(it);
  }
  if (t->status == KW_TICKET_DONE) {
    t->out.copy_to(out);
    kw_set_err(err, errlen, "");
    return KW_OK;
  }
  kw_set_err(err, errlen, t->msg.c_str());
  return t->status == KW_TICKET_CANCELLED ? KW_ESTATE : t->error;
}

int kw_cancel(kw_engine* engine, int64_t ticket, char* err, size_t errlen) {
  return kw_guarded(err, errlen, [&]() {
    if (!engine) kw::fail(KW_EINVAL, "NULL engine");
    kw::Engine& E = engine->e;
    std::lock_guard<std::mutex> lk(E.mu);
    auto it = E.tickets.find(ticket);
    if (it == E.tickets.end()) kw::fail(KW_EINVAL, "unknown ticket");
    kw::Ticket* t = it->second.get();
    if (kw::terminal(t->status)) return;
    t->cancel = true;
    if (t->status == KW_TICKET_QUEUED) {
      auto p = std::find(E.pending.begin(), E.pending.end(), t);
      if (p != E.pending.end()) E.pending.erase(p);
      E.finish(t, KW_TICKET_CANCELLED, KW_ESTATE, "cancelled");
    }
    E.cv_step.notify_one();
  });
}

int kw_engine_get_stats(kw_engine* engine, kw_engine_stats* st, char* err, size_t errlen) {
  return kw_guarded(err, errlen, [&]() {
    if (!engine || !st) kw::fail(KW_EINVAL, "NULL argument");
    kw::Engine& E = engine->e;
    std::lock_guard<std::mutex> lk(E.mu);
    st->submitted = E.n_submitted;
    st->completed = E.n_completed;
    st->failed = E.n_failed;
    st->cancelled = E.n_cancelled;
    st->queued = (int64_t)E.pending.size();
    st->running = E.n_running;
    st->device_bytes = E.held;
    st->device_budget = E.budget;
    st->cell_steps = E.cell_steps;
    st->busy_s = E.busy_s;
    st->launches = E.launches;
  });
}

int kw_engine_destroy(kw_engine* engine) {
  if (!engine) return KW_OK;
  kw::Engine& E = engine->e;
  {
    std::lock_guard<std::mutex> lk(E.mu);
    E.stop = true;
    for (kw::Ticket* t : E.pending) E.finish(t, KW_TICKET_CANCELLED, KW_ESTATE, "engine destroyed");
    E.pending.clear();
  }
  E.cv_admit.notify_all();
  E.cv_step.notify_all();
  for (std::thread& th : E.admit_thr)
    if (th.joinable()) th.join();
  if (E.step_thr.joinable()) E.step_thr.join();
  for (void* st : E.streams) kw::dev_stream_destroy(st);
  kw::dev_stream_destroy(E.admit_stream);
  kw::dev_host_free(E.flags);
  delete engine;
  return KW_OK;
}

}  // extern "C"
