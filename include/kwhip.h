/*
 * kwhip: MI355X (gfx950) implementation of the time-stepping loop of k-Wave's
 * kspaceFirstOrder2D / kspaceFirstOrder3D, behind a C ABI.
 *
 * The problem description mirrors the variable set that the reference hands
 * to its native engines after its MATLAB precompute
 * (+kwave/+legacy/private/kspaceFirstOrder_saveToDisk.m:46-108,253-380 and
 * +kwave/+legacy/kspaceFirstOrder3DC.m:468-546): grid sizes and spacings,
 * time step, medium maps, PML sizes and alphas, scaled source signals with
 * 1-based column-major index lists, the smoothed initial pressure, the
 * sensor index list and the record flags.  As in the reference's native
 * route the engine builds the PML profiles, the k-space correction kappa,
 * the staggered shift-derivative operators and the absorption operators
 * itself (in double precision, rounded once to single).
 *
 * All arrays are caller-owned, read-only, host memory in MATLAB column-major
 * order (x fastest).  Return codes: 0 on success, a negative KW_E* code on
 * failure with a message in `err` (truncated to errlen).  Every function is
 * synchronous unless stated; a kw_sim handle is not thread-safe.
 */
#ifndef KWHIP_H
#define KWHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KW_ABI_VERSION 2

/* error codes */
#define KW_OK 0
#define KW_EINVAL (-1)   /* invalid problem description */
#define KW_EUNSUP (-2)   /* valid in the reference but not supported here */
#define KW_EDEVICE (-3)  /* HIP runtime error */
#define KW_ENOMEM (-4)   /* device allocation failed */
#define KW_ESTATE (-5)   /* call not valid in the handle's current state */
#define KW_ENUMERIC (-6) /* engine: the simulation's pressure became non-finite
                            (an unstable time step); only that request fails */

/* sensor.record flags (reference record names,
 * +kwave/+legacy/private/kspaceFirstOrder_inputChecking.m:399-402) */
#define KW_REC_P (1u << 0)
#define KW_REC_P_MAX (1u << 1)
#define KW_REC_P_MIN (1u << 2)
#define KW_REC_P_RMS (1u << 3)
#define KW_REC_P_MAX_ALL (1u << 4)
#define KW_REC_P_MIN_ALL (1u << 5)
#define KW_REC_P_FINAL (1u << 6)
#define KW_REC_U (1u << 7)
#define KW_REC_U_MAX (1u << 8)
#define KW_REC_U_MIN (1u << 9)
#define KW_REC_U_RMS (1u << 10)
#define KW_REC_U_MAX_ALL (1u << 11)
#define KW_REC_U_MIN_ALL (1u << 12)
#define KW_REC_U_FINAL (1u << 13)
/* particle velocity shifted onto the non-staggered grid, time series at the
 * sensor points (kspaceFirstOrder_extractSensorData.m:36-57,111-122); the
 * intensities 'I' and 'I_avg' (kspaceFirstOrder_saveIntensity.m) are host
 * post-processing of it and of p */
#define KW_REC_U_NON_STAGGERED (1u << 14)
/* compressional and shear parts of the particle velocity at the sensor
 * points, time series (kspaceFirstOrder_extractSensorData.m:124-292; binary
 * sensor masks only, kspaceFirstOrder_inputChecking.m:1363-1366) */
#define KW_REC_U_SPLIT_FIELD (1u << 15)
/* the summed split density sum_d rho_d at the sensor points, time series
 * (the new toolbox's AcousticSensor densitySensor 'on',
 * +kwave/+toolbox/@AcousticSensor/AcousticSensor.m:153-173 with
 * @Sensor/Sensor.m ProcessSensorData 'sum') */
#define KW_REC_RHO (1u << 16)

/* directivity patterns (sensor.directivity_pattern) */
#define KW_DIR_PRESSURE 0
#define KW_DIR_GRADIENT 1

/* source modes, numbered as the reference's p_source_mode / u_source_mode
 * (kspaceFirstOrder_saveToDisk.m:225-231,262-268) */
#define KW_SRC_DIRICHLET 0
#define KW_SRC_ADDITIVE_NO_CORRECTION 1
#define KW_SRC_ADDITIVE 2

/* A time-varying source (the reference's source.p / source.ux ... with
 * p_mask / u_mask), signals already scaled by the caller as in
 * kspaceFirstOrder_scaleSourceTerms.m. */
typedef struct kw_time_source {
  int64_t flag;               /* number of time points; 0 = no source */
  int32_t mode;               /* KW_SRC_* */
  int64_t num_points;         /* source grid points */
  const uint64_t* index;      /* [num_points], 1-based column-major (find(mask)) */
  int64_t num_series;         /* rows of `input`: 1, num_points, or the label count */
  const uint64_t* sig_index;  /* [num_points] 1-based signal row per point for
                                 labelled masks, NULL otherwise */
  const double* input;        /* [num_series x flag], column-major */
} kw_time_source;

typedef struct kw_problem {
  int32_t dim;               /* 1, 2 or 3 (1: kspaceFirstOrder1D.m) */
  int64_t Nx, Ny, Nz;        /* Nz = 1 in 2D; Ny = Nz = 1 in 1D */
  int64_t Nt;                /* number of time steps */
  double dx, dy, dz, dt;
  double c_ref;              /* reference sound speed (setSoundSpeedRef.m) */

  /* medium: grids of Nx*Ny*Nz values, or NULL to use the scalar */
  const double* c0;
  double c0_scalar;
  const double* rho0;
  double rho0_scalar;
  const double* rho0_sg[3];  /* density at x+dx/2, y+dy/2, z+dz/2 (already
                                interpolated as kspaceFirstOrder3D.m:585-607);
                                NULL uses rho0 */
  int32_t nonlinear_flag;
  const double* BonA;
  double BonA_scalar;
  int32_t absorbing_flag;    /* 0 lossless, 1 power law, 2 stokes */
  const double* alpha_coeff; /* dB/(MHz^y cm), as medium.alpha_coeff */
  double alpha_coeff_scalar;
  double alpha_power;
  int32_t alpha_mode;        /* 0 both terms, 1 'no_absorption', 2 'no_dispersion' */
  /* medium.alpha_sign (kspaceFirstOrder_createAbsorptionVariables.m:78-81):
   * with alpha_sign_set = 1, tau is multiplied by sign(alpha_sign[0]) and
   * eta by sign(alpha_sign[1]), MATLAB's sign (-1, 0 or +1, so a 0 removes
   * the term); with alpha_sign_set = 0 both multipliers are +1 */
  double alpha_sign[2];
  int32_t alpha_sign_set;

  int64_t pml_size[3];
  double pml_alpha[3];
  int32_t use_kspace;        /* 1 unless 'UsekSpace' false */
  int32_t use_sg;            /* 1 unless 'UseSG' false */

  /* sources */
  const double* p0;          /* initial pressure (already smoothed) or NULL */
  kw_time_source p_source;
  kw_time_source u_source[3];

  /* sensor: binary mask as an index list; num_points = Nx*Ny*Nz with
   * index = NULL means every grid point in natural order */
  int64_t sensor_num_points;
  const uint64_t* sensor_mask_index; /* 1-based column-major */
  uint32_t record_flags;     /* KW_REC_* */
  int64_t record_start_index;/* 1-based */
  /* time reversal: [sensor_num_points x Nt] in the original (not reversed)
   * time order, or NULL */
  const double* time_reversal_boundary_data;
  /* Cartesian sensor points with 'CartInterp' 'linear'
   * (kspaceFirstOrder_createStorageVariables.m:376-398,
   * kspaceFirstOrder_extractSensorData.m:301-457): for each of the
   * sensor_num_points points the dim+1 grid points of the simplex that holds
   * it (record.tri: 1-based column-major grid indices) and its barycentric
   * coordinates there (record.bc), both [sensor_num_points x (dim+1)]
   * column-major; a sensor value is sum(f(tri) .* bc, 2).  sensor_mask_index
   * must then be NULL.  NULL for binary masks; not valid with time
   * reversal or KW_REC_U_SPLIT_FIELD. */
  const uint64_t* sensor_tri;
  const double* sensor_bc;
  /* 2-D sensor directivity (sensor.directivity_angle, _pattern, _size;
   * kspaceFirstOrder2D.m:945-947, directionalResponse.m): the angle of each
   * sensor point in radians [sensor_num_points], or NULL for none.  The
   * recorded p time series is then, per unique angle, the pressure filtered
   * in k-space by the pattern (KW_DIR_*) of element size directivity_size;
   * the statistics stay those of p.  2D binary masks only. */
  const double* directivity_angle;
  int32_t directivity_pattern;
  double directivity_size;
  /* medium.alpha_filter: a grid of Nx*Ny*Nz values in centred (fftshifted)
   * k-space order multiplying both power-law absorption operators
   * (kspaceFirstOrder_createAbsorptionVariables.m:56-76), or NULL */
  const double* alpha_filter;
  /* cuboid-corners sensor (the native contract's sensor_mask_corners,
   * kspaceFirstOrder_saveToDisk.m, kspaceFirstOrder_inputChecking.m:
   * 1610-1639): sensor_num_cuboids columns of [x1 y1 z1 x2 y2 z2] (2*dim
   * rows, 1-based, inclusive), column-major; the sensor points are the
   * cuboids' grid points in column-major order, cuboid after cuboid, and
   * sensor_num_points must equal their total.  sensor_mask_index must then
   * be NULL. */
  const uint64_t* sensor_mask_corners;
  int64_t sensor_num_cuboids;
  /* kspaceFirstOrderAS (+kwave/+legacy/kspaceFirstOrderAS.m) with the
   * 'WSWA-FFT' radial symmetry: 1 when this 2-D problem is the radially
   * expanded grid of an axisymmetric one (x axial, y radial, Ny four times
   * the radial size, every input laid out as kspaceFirstOrderAS.m:703-733
   * lays out the fields: p-like grids [f, 0, -flip(f(2:end)), -f, 0,
   * flip(f(2:end))] along y, the y-staggered ones [f, flip(f), -f,
   * -flip(f)], material grids with the same layouts unsigned); pml_size[1]
   * is the outer radial PML of the radial grid.  The library then builds
   * the y PML on every copy (getPML.m with axisymmetric true) and adds the
   * u_r / r term of the divergence (:954-969).  Lossless or Stokes media
   * (with or without B/A), p0, pressure and axial velocity sources and
   * time-reversal data given on the expanded grid with the signs of the
   * p-like layout, radial velocity sources with those of the y-staggered
   * layout, and binary or interpolated (sensor_tri on the first copy)
   * sensors.  2: the radial grid
   * itself (Ny the radial size, inputs as kspaceFirstOrderAS.m prepares
   * them), the layout of the file contract: kw_save_problem writes it with
   * axisymmetric_flag = 1 and kw_run_file lays such a file out on the
   * expanded grid and writes the radial grid's outputs; kw_run refuses 2.
   * 3: the expanded grid of 'WSWS-FFT' (Ny = 2 x radial size - 2, p-like
   * grids [f, flip(f(2:end-1))], y-staggered ones [f(1:end-1),
   * flip(f(1:end-1))], kspaceFirstOrderAS.m:720-722, 966-977; sources and
   * time-reversal points at columns j and 2Ny_r - 2 - j, both +; radial
   * velocity sources at j and 2Ny_r - 3 - j, + and -, 'dirichlet' and
   * 'additive-no-correction' ones with no point on row Ny/2 (KW_EUNSUP
   * otherwise)).  0 otherwise. */
  int32_t axisymmetric;
} kw_problem;

typedef struct kw_sizes {
  int64_t num_sensor_points;
  int64_t num_recorded_time_points;
  int64_t grid_cells;
} kw_sizes;

/* Caller-allocated outputs; a NULL pointer skips that output.  Shapes:
 * time series [num_sensor_points x num_recorded_time_points] column-major,
 * statistics [num_sensor_points], *_all and *_final [grid_cells]. */
typedef struct kw_outputs {
  float* p;
  double* p_max;
  double* p_min;
  double* p_rms;
  double* p_max_all;
  double* p_min_all;
  float* p_final;
  float* u[3];
  double* u_max[3];
  double* u_min[3];
  double* u_rms[3];
  double* u_max_all[3];
  double* u_min_all[3];
  float* u_final[3];
  float* u_non_staggered[3]; /* [points x recorded steps] per component */
  float* u_split_p[3];       /* [points x recorded steps] per component */
  float* u_split_s[3];
  float* rho;                /* KW_REC_RHO: [points x recorded steps] */
} kw_outputs;

typedef struct kw_sim kw_sim;

/* ABI version and the compute backend of this library build. */
int kw_abi_version(void);
const char* kw_backend(void);

/* Output sizes of a problem.
 * Replaces: kspaceFirstOrder_createStorageVariables.m:57-400 (allocation). */
int kw_output_sizes(const kw_problem* prob, kw_sizes* sizes, char* err, size_t errlen);

/* Run a whole simulation and copy its outputs.
 * Replaces: the hot loop kspaceFirstOrder3D.m:769-1208 (2D: :712-1044)
 * plus the native-engine call of kspaceFirstOrder3DC.m:546. */
int kw_run(const kw_problem* prob, kw_outputs* out, char* err, size_t errlen);

/* Device-resident simulation: inputs are copied to the device and all
 * operators built by kw_sim_create; kw_sim_step advances nsteps time steps
 * asynchronously on the handle's stream; kw_sim_read waits and copies the
 * outputs recorded so far. */
int kw_sim_create(const kw_problem* prob, kw_sim** sim, char* err, size_t errlen);
int kw_sim_step(kw_sim* sim, int64_t nsteps, char* err, size_t errlen);
int kw_sim_sync(kw_sim* sim, char* err, size_t errlen);
int kw_sim_read(kw_sim* sim, kw_outputs* out, char* err, size_t errlen);
int64_t kw_sim_steps_done(const kw_sim* sim);
int64_t kw_sim_device_bytes(const kw_sim* sim);
int kw_sim_destroy(kw_sim* sim);

/* Entry points of the new toolbox's time-domain solver, whose run() may be
 * called repeatedly on the same state, with a new time step each time
 * (+kwave/+toolbox/@TimeDomainSolver/run.m:58-101,
 * @AcousticSolver/executeTimeStep.m:35-260):
 *
 * kw_sim_extend: grow the simulation to Nt_total time steps (at least its
 * current number); recorded time series keep their columns and gain the
 * new ones (@AcousticSensor/AcousticSensor.m initialiseSensorData appends
 * columns).  Not for time reversal or the no-allocation entry point.
 *
 * kw_sim_change_dt: the next step is the transition from time step dt_prev
 * to dt_new (executeTimeStep.m:45-55,143-176): its velocity update uses the
 * mean step (dt_prev + dt_new)/2 (in single precision, as the reference's
 * single prevTimeStep makes it), the PML of that step, the two-step k-space
 * correction kappa and the split correction term dt * kappaSplit(pml(u))
 * (setkSpaceCorrection.m:36-57, kappaSplitCorrection.m:56-60); its density
 * update and every later step use dt_new.  dt_prev = 0 gives the first step
 * after an initial particle velocity.  A dt_new equal to dt_prev in single
 * precision changes nothing.
 *
 * kw_sim_init_velocity: the first step as the initial condition with a
 * given particle velocity (executeTimeStep.m:128-142): p = p0, rho_d =
 * p0 / (dim c0^2), u_d = u[d] (already on the staggered grid; Nx*Ny*Nz
 * values each, column-major), recorded like any first step.  Needs p0 and
 * no step taken.
 *
 * kw_sim_get_fields: copy the current p, u_d and rho_d (the solver's
 * pressure, velocity and densitySplit) to host arrays of Nx*Ny*Nz floats;
 * NULL skips a field. */
int kw_sim_extend(kw_sim* sim, int64_t Nt_total, char* err, size_t errlen);
int kw_sim_change_dt(kw_sim* sim, double dt_prev, double dt_new, char* err, size_t errlen);
int kw_sim_init_velocity(kw_sim* sim, const double* const u[3], char* err, size_t errlen);
int kw_sim_get_fields(kw_sim* sim, float* p, float* const u[3], float* const rho[3], char* err,
                      size_t errlen);

/* Per-kernel-kind launch statistics of a timed run: summed device time
 * (each launch bracketed by events), launch count, and the compulsory HBM
 * bytes of those launches (every array a launch reads or writes, counted
 * once per launch). */
#define KW_NKINDS 9
typedef struct kw_kernel_stats {
  double ms[KW_NKINDS];
  int64_t launches[KW_NKINDS];
  double bytes[KW_NKINDS];
} kw_kernel_stats;

/* Advance nsteps steps timed with device events on the handle's stream;
 * *ms receives the elapsed device time.  With stats non-NULL every launch
 * is additionally bracketed by events (slower; use a separate run for the
 * headline time). */
int kw_sim_step_timed(kw_sim* sim, int64_t nsteps, double* ms, kw_kernel_stats* stats,
                      char* err, size_t errlen);
/* Pencils per block the library launches for a problem's grid, for the x,
 * strided y and last-axis pass families (tiles[1] is 0 in 1-D and 2-D).
 * They come from a tile autotuner: the first time a grid shape and equation
 * of state (absorbing, nonlinear) is seen, a scratch simulation of that
 * shape steps under each candidate with every launch timed by device
 * events, and each family takes its fastest candidate.  Choices are cached
 * per process and, when the environment variable KW_TUNE_FILE names a text
 * file, in that file for later processes; grids of 2^24 cells and more, the
 * no-allocation entry point (which only reads the cache) and KW_TUNE=0 keep
 * the built-in tiles.  Tiles never change results.  No counterpart in the
 * reference, whose native engines fix their launch shapes. */
int kw_tiles(const kw_problem* prob, int32_t tiles[3], char* err, size_t errlen);

/* Name of kernel kind k (0 <= k < KW_NKINDS). */
const char* kw_kind_name(int k);

/* ---------------------------------------------------------------------------
 * Engine: asynchronous submission of many simulations, stepped together.
 *
 * Replaces the one-simulation-per-call contract of kspaceFirstOrder3D.m /
 * kspaceFirstOrder3DC.m:441-546 for callers that run many simulations
 * (batches of forward runs and time-reversal reconstructions, request
 * streams).  kw_submit copies the problem (every array) into the engine
 * before it returns, so the caller's buffers may be freed at once.
 * Admission threads move queued simulations to the device, in submission
 * order, while their estimated device footprint fits the engine's budget
 * (several are built at once: the host-side operator construction of a
 * large grid takes a large part of a second); a stepping thread
 * advances all admitted simulations together, one launch per pass for all
 * simulations with the same dimension and equation of state (and the same
 * grid size where fixed-length kernels exist for it), whatever their media,
 * sources, sensors and current step.  Groups run on a pool of streams and
 * each group's rounds of steps are captured once as a graph and replayed.
 * Axisymmetric problems (kw_problem.axisymmetric 1 or 3) group with their
 * own kind and size and step one by one on their group's stream.
 * Results are the same as kw_run on the same problem (the per-simulation
 * arithmetic does not change).  Every engine function is thread-safe.
 * ------------------------------------------------------------------------- */
typedef struct kw_engine kw_engine;

typedef struct kw_engine_config {
  int64_t device_bytes;    /* device memory the engine may hold for simulations;
                              0: 90% of the free memory at creation */
  int32_t steps_per_round; /* steps between admission/retirement decisions; 0: 32 */
  int32_t admit_threads;   /* simulations built on the host and device at the
                              same time (operators, uploads, first step); 0: 3 */
} kw_engine_config;

#define KW_TICKET_QUEUED 0
#define KW_TICKET_RUNNING 1
#define KW_TICKET_DONE 2
#define KW_TICKET_FAILED 3
#define KW_TICKET_CANCELLED 4

typedef struct kw_ticket_info {
  int32_t status;          /* KW_TICKET_* */
  int32_t error;           /* KW_E* code when FAILED */
  double submit_s;         /* engine clock (seconds since creation) */
  double start_s;          /* admitted to the device (-1 before) */
  double done_s;           /* finished, failed or cancelled (-1 before) */
  int64_t steps_done;
  int64_t steps_total;
  int64_t grid_cells;
  int64_t device_bytes;    /* held on the device while running */
} kw_ticket_info;

typedef struct kw_engine_stats {
  int64_t submitted, completed, failed, cancelled, queued, running;
  int64_t device_bytes;    /* held now */
  int64_t device_budget;
  double cell_steps;       /* grid cells x steps executed */
  double busy_s;           /* time with at least one simulation running */
  int64_t launches;        /* kernel launches issued by the stepping thread */
} kw_engine_stats;

int kw_engine_create(const kw_engine_config* cfg, kw_engine** engine, char* err, size_t errlen);
int kw_submit(kw_engine* engine, const kw_problem* prob, int64_t* ticket, char* err,
              size_t errlen);
int kw_poll(kw_engine* engine, int64_t ticket, kw_ticket_info* info, char* err, size_t errlen);
/* Blocks until the ticket has finished, copies its outputs (as kw_run) into
 * `out` (NULL: discard them) and releases the ticket.  Returns the
 * simulation's error code if it failed, KW_ESTATE if it was cancelled.
 * A simulation whose pressure becomes non-finite (NaN or Inf, flagged on
 * the device as the pressure is formed and checked after every round of
 * steps) fails alone with KW_ENUMERIC and leaves its group; the other
 * simulations run on, unaffected.  A HIP error while a group's round is
 * being launched fails the members of that group. */
int kw_wait(kw_engine* engine, int64_t ticket, kw_outputs* out, char* err, size_t errlen);
int kw_cancel(kw_engine* engine, int64_t ticket, char* err, size_t errlen);
int kw_engine_get_stats(kw_engine* engine, kw_engine_stats* stats, char* err, size_t errlen);
/* Cancels queued and running simulations, stops the threads, frees all. */
int kw_engine_destroy(kw_engine* engine);

/* kw_run without any device allocation.  It has no counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * running while the surrounding application captures a HIP graph, serving
 * loops and harnesses that forbid hipMalloc).  Every device buffer of the
 * simulation (fields, spectra, operators, the staged inputs and recorded
 * outputs) is carved out of a static scratch array of KW_NOALLOC_MAX_BYTES
 * bytes that each device holds from the library's load until the process
 * ends; a problem whose device footprint (kw_sim_device_bytes of the same
 * problem, plus 256-byte alignment per buffer) exceeds it fails with
 * KW_ENOMEM.  Calls in one process are serialised by a lock in the
 * library, since they share the array; the array keeps the last call's
 * contents, and each call clears what it needs.  Runs on the null stream
 * of the current device.  Outputs are those of kw_run. */
#define KW_NOALLOC_MAX_BYTES (128ll << 20)
int kw_run_noalloc(const kw_problem* prob, kw_outputs* out, char* err, size_t errlen);

/* The native engines' file contract.
 *
 * kw_save_problem writes the HDF5 input file of a problem: the variables,
 * names and storage of +kwave/+legacy/private/kspaceFirstOrder_saveToDisk.m
 * (uint64 sizes, flags and 1-based index lists; single-precision reals;
 * MATLAB array sizes; writeMatrix.m's domain_type / data_type attributes;
 * writeAttributes.m's file attributes, version 1.2, file_type 'input').
 * Replaces the 'SaveToDisk' branch of kspaceFirstOrder3D.m:668-686.  2-D
 * and 3-D forward problems with binary or cuboid-corner sensor masks only,
 * as the contract (KW_EUNSUP otherwise).
 *
 * kw_run_file runs such an input file (written by the reference's MATLAB
 * code or by kw_save_problem) and writes the output file the reference's
 * MATLAB side reads back (kspaceFirstOrder3DC.m:571-777): /p, /p_max, /p_min,
 * /p_rms, /p_max_all, /p_min_all, /p_final, /ux ... /uz_non_staggered in
 * single precision with MATLAB array sizes, per-cuboid groups (/p/1, ...) for
 * cuboid-corner sensors, and /Nx, /Ny, /Nz, /Nt, /pml_*_size.  record_flags
 * are KW_REC_* (the command-line flags --p_raw, --p_max, ...), and
 * record_start_index is the -s option.  Replaces the native binary call of
 * kspaceFirstOrder3DC.m:546 (the executable kspaceFirstOrder-HIP wraps it).
 * Errors in the file (missing or malformed variables, unsupported HDF5
 * features or simulation types) return KW_EINVAL or KW_EUNSUP. */
int kw_save_problem(const kw_problem* prob, const char* path, char* err, size_t errlen);
int kw_run_file(const char* input_file, const char* output_file, uint32_t record_flags,
                int64_t record_start_index, char* err, size_t errlen);

/* Smoothing filter: out = real(ifftn(fftn(in) .* ifftshift(win))) on a 1D,
 * 2D or 3D grid, computed on the device.  `win` is the centred window of
 * getWin (not shifted).  Replaces the transform of +kwave/+legacy/smooth.m;
 * the caller restores the maximum. */
int kw_smooth(int32_t dim, const int64_t N[3], const double* in, const double* win,
              double* out, char* err, size_t errlen);

#ifdef __cplusplus
}
#endif

#endif /* KWHIP_H */
