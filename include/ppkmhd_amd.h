/*
 * ppkmhd_amd.h -- C ABI of the MI355X (gfx950) MUSCL-Hancock + HLLD + constrained-transport
 * MHD time step.
 *
 * Every entry point is extern "C", takes plain pointers and sizes, never throws, and reports
 * failure through an integer status (PPK_OK == 0) plus a per-handle message
 * (ppk_mhd_last_error).  One handle owns one simulation: its device buffers, its HIP stream
 * and its time-loop counters.  A handle is not thread-safe; calls on one handle are
 * stream-ordered.
 *
 * Array layout used across this ABI ("reference layout"): 8 FP64 variables per cell in the
 * order ID=0 (density), IP=1 (total energy), IU=2, IV=3, IW=4 (momentum), IA=5, IB=6, IC=7
 * (magnetic field), stored variable-major with i fastest:
 *     U[((v*ksize + k)*jsize + j)*isize + i]
 * over the ghost-padded box isize = nx + 2*gw (same for j, k; ksize = 1 in 2D), gw = 3.
 * IA/IB/IC hold the face-centred field on the LEFT face of the cell (x-face i-1/2, y-face
 * j-1/2, z-face k-1/2); in 2D IC is a cell-centred Bz.  This is the LayoutLeft View
 * DataArray{2d,3d}(isize, jsize[, ksize], nbvar) of the reference
 * (src/shared/kokkos_shared.h:22-42, var order src/shared/enums.h:17-34).
 */
#ifndef PPKMHD_AMD_H_
#define PPKMHD_AMD_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------------------------ */
enum ppk_status {
  PPK_OK = 0,
  PPK_ERR_INVALID_ARGUMENT = 1, /* bad pointer, size, enum value or unsupported option   */
  PPK_ERR_HIP = 2,              /* a HIP runtime call failed (message names it)           */
  PPK_ERR_OUT_OF_MEMORY = 3,
  PPK_ERR_UNSUPPORTED = 4,      /* option the reference does not define for this path    */
  PPK_ERR_CONFIG = 5            /* .ini file missing / unreadable / unknown solver name   */
};

/* ---- enums, same integer values as the reference ----------------------------------------- */
/* src/shared/enums.h:141-148 (RiemannSolverType) */
enum ppk_riemann { PPK_RIEMANN_APPROX = 0, PPK_RIEMANN_LLF = 1, PPK_RIEMANN_HLL = 2,
                   PPK_RIEMANN_HLLC = 3, PPK_RIEMANN_HLLD = 4 };
/* src/shared/enums.h:152-159 (BoundaryConditionType) */
enum ppk_bc { PPK_BC_UNDEFINED = 0, PPK_BC_DIRICHLET = 1, PPK_BC_NEUMANN = 2,
              PPK_BC_PERIODIC = 3 };
/* arithmetic of the step.  PPK_FP64: double, the compiler may contract a*b+c into FMA.
 * PPK_FP64_STRICT: double with every product and sum rounded separately, in the reference's
 * operand order (bit-identical to the reference's double arithmetic).  PPK_FP32_BULK: trace
 * and Riemann solves in float; state, CT/EMF update and dt stay double. */
enum ppk_precision { PPK_FP64 = 0, PPK_FP32_BULK = 1, PPK_FP64_STRICT = 2 };
/* kernel organisation of the step */
enum ppk_kernel_path {
  PPK_PATH_AUTO = 0,   /* fused kernels where available, else staged                         */
  PPK_PATH_STAGED = 1, /* one kernel per reference stage, intermediates in HBM                */
  PPK_PATH_FUSED = 2,  /* fused kernels: in 3D FP64 the single-pass step (state read once and
                        * written once), else the three-kernel form                          */
  PPK_PATH_FUSED_RECORDS = 3 /* the three-kernel fused form in every case (half-step records,
                        * face/edge solves, update; records and fluxes in HBM)               */
};
/* dt semantics of ppk_mhd_step */
enum ppk_dt_mode {
  PPK_DT_FIXED = 0, /* use the dt argument (parity runs at fixed dt), clamped to tEnd - t    */
  PPK_DT_CFL = 1    /* reference semantics: dt = cfl / max(sum_d (c_f,d+|v_d|)/dx_d) from the
                       current state, clamped to tEnd - t (src/shared/SolverBase.cpp:149-179) */
};
/* initial conditions (src/muscl/SolverMHDMuscl.h:649-713 problem names) */
enum ppk_problem_kind { PPK_PROBLEM_ORSZAG_TANG = 0, PPK_PROBLEM_BLAST = 1,
                        PPK_PROBLEM_KELVIN_HELMHOLTZ = 2, PPK_PROBLEM_WAVE = 3,
                        PPK_PROBLEM_IMPLODE = 4,
                        /* hydro solvers only (src/muscl/SolverHydroMuscl.h:435-520) */
                        PPK_PROBLEM_GRESHO_VORTEX = 5, PPK_PROBLEM_ISENTROPIC_VORTEX = 6,
                        PPK_PROBLEM_FOUR_QUADRANT = 7 /* 2D */,
                        /* with static gravity ([gravity] static=yes) */
                        PPK_PROBLEM_RAYLEIGH_TAYLOR = 8, PPK_PROBLEM_RISING_BUBBLE = 9,
                        PPK_PROBLEM_DISK = 10 /* point-mass gravity */,
                        /* MHD solvers only (src/muscl/SolverMHDMuscl.h:686-695) */
                        PPK_PROBLEM_ROTOR = 11, PPK_PROBLEM_FIELD_LOOP = 12 };

/*
 * Solver parameters.  Mirrors the fields of HydroParams / HydroSettings that the MHD MUSCL
 * path reads (src/shared/HydroParams.h:28-234); derived fields (dx, smallp, smallpp, gamma6)
 * follow HydroParams::init (src/shared/HydroParams.cpp:421-453).  ppk_mhd_params_from_ini
 * fills it exactly as HydroParams::setup does, including the float rounding of .ini reals.
 */
typedef struct ppk_mhd_params {
  int32_t dim;           /* 2 or 3                                                     */
  int32_t nx, ny, nz;    /* interior cells (nz = 1 in 2D)                              */
  int32_t ghost_width;   /* 3 for MHD (src/shared/HydroParams.cpp:56-71)               */
  int32_t riemann;       /* enum ppk_riemann: HLLD, HLL or LLF                         */
  int32_t bc[6];         /* enum ppk_bc for xmin, xmax, ymin, ymax, zmin, zmax         */
  int32_t n_step_max;    /* [run] nStepmax                                             */
  int32_t precision;     /* enum ppk_precision                                         */
  int32_t kernel_path;   /* enum ppk_kernel_path                                       */
  /* [OTHER] implementationVersion (src/shared/HydroParams.cpp:446-451): 0 or 1 (both run
   * the version-0 stages: version 1 only accumulates the same fluxes with atomics); 2 (MHD's
   * updated-primitive scheme, SolverMHDMuscl.cpp:430-456) is refused with PPK_ERR_UNSUPPORTED */
  int32_t implementation_version;
  double xmin, xmax, ymin, ymax, zmin, zmax;
  double dx, dy, dz;
  double gamma0, gamma6, cfl, slope_type, smallr, smallc, smallp, smallpp, cIso;
  double tEnd;
  /* hydro solvers: [gravity] static or point (src/shared/SolverBase.cpp:138-141): 1 adds the
   * gravity predictor to the face states, the gravity source term after the update and the
   * gravity-corrected CFL dt, with the per-cell field the problem sets (uniform for
   * rayleigh_taylor / rising_bubble, the point mass for disk, 0 for problems that set none).
   * The MHD solvers ignore it, as the reference's do. */
  int32_t gravity_static;
  int32_t reserved2;     /* must be 0                                                  */
} ppk_mhd_params;

/*
 * Initial-condition parameters of every problem: MHD (Orszag-Tang, blast, Kelvin-Helmholtz,
 * linear wave, implode, rotor, field loop) and hydro (implode, blast, Kelvin-Helmholtz,
 * Gresho vortex, isentropic vortex, four quadrant, Rayleigh-Taylor, rising bubble, disk).
 * Field defaults are the reference's (the ...Params.h files of src/shared/problems).  Extensions beyond
 * the reference are marked "ext" and default to the reference behaviour.
 */
typedef struct ppk_mhd_problem {
  int32_t kind;          /* enum ppk_problem_kind                                      */
  int32_t kh_mode;       /* KH perturbation mode                                        */
  int32_t kh_p_sine;     /* KH perturbation_sine                                        */
  int32_t kh_p_sine_rob; /* KH perturbation_sine_robertson                              */
  int32_t kh_p_rand;     /* KH perturbation_rand (counter-based RNG, see below)         */
  int32_t kh_seed;       /* KH rand_seed                                                */
  int32_t kh_axis;       /* ext: shear-normal axis for the 3D layer (2 = z, reference; 1 = y) */
  int32_t wave_type;     /* linear wave 0 fast, 1 Alfven, 2 slow, 3 contact              */
  int32_t implode_shape; /* implode shape                                               */
  int32_t pad;
  double ot_kt;          /* OrszagTang kt (3D only)                                     */
  double ot_b0_scale;    /* ext: multiplies B0 (1.0 = reference)                        */
  double blast_radius, blast_center_x, blast_center_y, blast_center_z;
  double blast_density_in, blast_density_out, blast_pressure_in, blast_pressure_out;
  double blast_bx, blast_by, blast_bz; /* ext: uniform B (0.5, 0.5, 0.5 = reference)     */
  double kh_d_in, kh_d_out, kh_pressure, kh_vflow_in, kh_vflow_out, kh_amplitude;
  double kh_w0, kh_delta;
  double kh_bx, kh_by, kh_bz; /* ext: uniform B of the 3D random layer (0 = reference)  */
  double wave_amplitude;
  /* WaveParams re-reads the mesh extents and gamma0 with its own defaults (3.0, 1.5, 1.5,
   * 1.66667; src/shared/problems/WaveParams.h:33-45), so they are carried separately */
  double wave_gamma0, wave_lx, wave_ly, wave_lz;
  /* implode inner / outer state: rho, p, u, v, w, Bx, By, Bz
   * (src/shared/problems/ImplodeParams.h:40-58) */
  double implode_in[8], implode_out[8];
  /* Gresho vortex rho0, Ma (src/shared/problems/GreshoParams.h) */
  double gresho_rho0, gresho_ma;
  /* isentropic vortex: ambient density, pressure, temperature, velocity, centre, strength
   * (src/shared/problems/IsentropicVortexParams.h) */
  double iv_rho_a, iv_p_a, iv_t_a, iv_u_a, iv_v_a, iv_w_a, iv_x, iv_y, iv_z, iv_beta;
  /* four-quadrant 2D Riemann problem: [riemann2d] config_number (0-based Lax-Liu
   * configuration), quadrant corner x, y (src/muscl/SolverHydroMuscl.cpp:72-89) */
  int fq_config, fq_reserved;
  double fq_x, fq_y;
  /* Rayleigh-Taylor: perturbation amplitude, densities below / above, uniform gravity
   * (src/shared/problems/RayleighTaylorInstabilityParams.h) */
  double rti_amplitude, rti_d0, rti_d1, rti_gx, rti_gy, rti_gz;
  /* rising bubble: radius, centre, densities in / out, uniform gravity
   * (src/shared/problems/RisingBubbleParams.h) */
  double rb_R, rb_x0, rb_y0, rb_z0, rb_din, rb_dout, rb_gx, rb_gy, rb_gz;
  /* disk around a point mass (src/shared/problems/DiskParams.h) and the softened point-mass
   * field it sets ([gravity] x, y, z, G*M, soften; src/shared/problems/PointSourceGravity.h) */
  double disk_radius, disk_radius_inner, disk_xc, disk_yc, disk_zc, disk_ref_density,
      disk_contrast_density, disk_contrast_width, disk_ref_sound_speed;
  double pg_x, pg_y, pg_z, pg_gm, pg_eps;
  /* MHD rotor: radii, spin speed, pressure, field (src/shared/problems/RotorParams.h) */
  double rotor_r0, rotor_r1, rotor_u0, rotor_p0, rotor_b0;
  /* MHD field loop: radius, density inside, vector-potential amplitude, flow speed
   * (src/shared/problems/FieldLoopParams.h) */
  double loop_radius, loop_density_in, loop_amplitude, loop_vflow;
} ppk_mhd_problem;

typedef struct ppk_mhd_solver ppk_mhd_t;

/* ---- library ----------------------------------------------------------------------------- */
int ppk_mhd_params_size(void);  /* sizeof(ppk_mhd_params), for FFI layout checks  */
int ppk_mhd_problem_size(void); /* sizeof(ppk_mhd_problem)                         */
const char *ppk_mhd_version(void);
/* number of HIP devices visible (0 when none); never fails */
int ppk_mhd_device_count(void);

/*
 * Parse an .ini file the way the reference does: case-insensitive sections and keys
 * (src/utils/config/inih/INIReader.cpp:100-105), reals through strtof, i.e. rounded to float
 * (src/utils/config/ConfigMap.cpp:37-46), HydroParams::setup defaults
 * (src/shared/HydroParams.cpp:28-217) and the problem parameter defaults.  solver_name
 * receives [run] solver_name (NUL-terminated, truncated to name_len).  Solver names
 * MHD_Muscl_2D/3D give ghost_width 3 (for ppk_mhd_create), Hydro_Muscl_2D/3D ghost_width 2
 * (for ppk_hydro_create); any other name fails with PPK_ERR_CONFIG.
 * Replaces: HydroParams::setup + broadcast_parameters (src/main.cpp:133-137).
 */
int ppk_mhd_params_from_ini(const char *ini_path, ppk_mhd_params *params,
                            ppk_mhd_problem *problem, char *solver_name, int name_len);
/* Same, from an in-memory .ini text. */
int ppk_mhd_params_from_ini_text(const char *ini_text, ppk_mhd_params *params,
                                 ppk_mhd_problem *problem, char *solver_name, int name_len);
/* Same, through a caller's key lookup instead of a file: lookup(ctx, section, name, value,
 * value_size) writes the raw value string of [section] name (NUL-terminated) and returns 1,
 * or returns 0 when the key is absent; the typed parsing and defaults are the ones above.
 * A reference-side binding passes the ConfigMap the solver was constructed with
 * (SolverFactory::create(name, HydroParams &, ConfigMap &), src/shared/SolverFactory.h:43),
 * so the problem parameters come from the same source as the reference's problem structs
 * (BlastParams(configMap), ..., src/shared/problems/). */
typedef int (*ppk_ini_lookup)(void *ctx, const char *section, const char *name, char *value,
                              int value_size);
int ppk_mhd_params_from_lookup(ppk_ini_lookup lookup, void *ctx, ppk_mhd_params *params,
                               ppk_mhd_problem *problem, char *solver_name, int name_len);
/* Fill the derived fields (dx, dy, dz, smallp, smallpp, gamma6) from the primary ones,
 * as HydroParams::init does (src/shared/HydroParams.cpp:421-453). */
int ppk_mhd_params_init(ppk_mhd_params *params);

/* ---- solver handle ----------------------------------------------------------------------- */
/*
 * Create a solver on the current HIP device with its own stream.  Allocates the state
 * (two ping-pong copies, as U / U2 of src/muscl/SolverMHDMuscl.h:73-77) and workspace.
 * Replaces: SolverMHDMuscl<dim>::SolverMHDMuscl allocation part
 * (src/muscl/SolverMHDMuscl.h:212-387) registered under MHD_Muscl_2D / MHD_Muscl_3D.
 */
int ppk_mhd_create(const ppk_mhd_params *params, ppk_mhd_t **handle);
void ppk_mhd_destroy(ppk_mhd_t *handle);
/* Message of the last failed call on this handle ("" when none).  handle may be NULL for
 * the last failure of ppk_mhd_create / params parsing on this thread. */
const char *ppk_mhd_last_error(const ppk_mhd_t *handle);

/* Initial condition computed by the library (src/muscl/MHDInitFunctors{2D,3D}.h), then
 * ghost fill, as SolverMHDMuscl ctor does (src/muscl/SolverMHDMuscl.h:389-402).
 * Resets t = params.tCurrent (0), iteration = 0. */
int ppk_mhd_init_problem(ppk_mhd_t *h, const ppk_mhd_problem *problem);
/* The same with the initial state computed on the GPU (one thread per cell; no host array,
 * no PCIe transfer).  The cells' sin / cos / exp come from the device math library, so a
 * state can differ from ppk_mhd_init_problem's in the last bits where those appear.
 * Replaces: the reference's init functors run on the device (src/muscl/MHDInitFunctors2D.h,
 * src/muscl/MHDInitFunctors3D.h). */
int ppk_mhd_init_problem_device(ppk_mhd_t *h, const ppk_mhd_problem *problem);
/* The initial state of a problem computed on the host into U (count = 8*isize*jsize*ksize),
 * no GPU needed: what ppk_mhd_init_problem uploads. */
int ppk_mhd_initial_state(const ppk_mhd_params *params, const ppk_mhd_problem *problem,
                          double *U, size_t count);

/* Host <-> device copies of the current state in the reference layout (count =
 * 8*isize*jsize*ksize doubles).  download synchronizes the handle's stream. */
int ppk_mhd_upload(ppk_mhd_t *h, const double *U, size_t count);
int ppk_mhd_download(ppk_mhd_t *h, double *U, size_t count);
/* Device pointer of the current state (same layout); valid until the next step. */
int ppk_mhd_device_state(ppk_mhd_t *h, double **dU);
/* The handle's HIP stream (hipStream_t), for callers that order their own work on it. */
int ppk_mhd_stream(ppk_mhd_t *h, void **stream);

/* Ghost fill of the current state, faces X then Y then Z over the full transverse range.
 * Replaces: make_boundaries (src/muscl/SolverMHDMuscl.cpp:22-64,
 * src/shared/SolverBase.cpp:339-537, src/shared/BoundariesFunctors.h:534-1057). */
int ppk_mhd_make_boundaries(ppk_mhd_t *h);

/* CFL time step of the current state: dt = cfl / invDt, NOT clamped to tEnd (the clamp is
 * part of the time loop).  Synchronizes.
 * Replaces: SolverMHDMuscl::compute_dt_local (src/muscl/SolverMHDMuscl.h:722-741). */
int ppk_mhd_compute_dt(ppk_mhd_t *h, double *dt);

/* No counterpart in the reference.  For callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops): ppk_mhd_compute_dt on a caller-owned device
 * state dU (this library's layout, 8*isize*jsize*ksize doubles) without a solver handle and
 * without any allocation.  Fills dU's ghosts in place, then writes dt = cfl / invDt (the
 * value ppk_mhd_compute_dt returns for the same state) to the device double *dt, all on
 * `stream` (a hipStream_t; NULL = the null stream); does not synchronize.  The kernels' control
 * block and CFL partial maxima live in one 8 MiB device array per GPU, kept until the process
 * ends; grids of more than 2^28 padded cells are refused with PPK_ERR_INVALID_ARGUMENT.  All
 * calls on one GPU share that array: calls that overlap on one GPU (other streams or threads)
 * race and must be serialised by the caller. */
int ppk_mhd_compute_dt_noalloc(const ppk_mhd_params *params, double *dU, double *dt,
                               void *stream);

/*
 * One time step (godunov_unsplit_impl, src/muscl/SolverMHDMuscl.cpp:373-572): ghost fill,
 * primitives, dt (per dt_mode), trace, fluxes, EMFs, flux update and CT update, written to
 * the other ping-pong buffer; then t += dt, iteration += 1
 * (src/shared/SolverBase.cpp:206-220).  Stream-ordered; does not synchronize.
 * Replaces: SolverBase::next_iteration -> SolverMHDMuscl::next_iteration_impl (the IO part
 * of next_iteration_impl is not on the path).
 */
int ppk_mhd_step(ppk_mhd_t *h, double dt, int dt_mode);
/* nsteps calls of ppk_mhd_step(h, dt, dt_mode) without host synchronization, stopping on
 * the device once t >= tEnd - 1e-14 or iteration >= n_step_max (SolverBase::finished,
 * src/shared/SolverBase.cpp:195-201); steps past that point do no work. */
int ppk_mhd_run(ppk_mhd_t *h, int nsteps, double dt, int dt_mode);
/* Current time, last dt and iteration counter (synchronizes). */
int ppk_mhd_get_time(ppk_mhd_t *h, double *t, double *dt, int64_t *iteration);
/* Set time and iteration (restart); iteration parity selects the ping-pong buffer. */
int ppk_mhd_set_time(ppk_mhd_t *h, double t, int64_t iteration);
/* 1 when SolverBase::finished() holds (synchronizes). */
int ppk_mhd_finished(ppk_mhd_t *h, int *finished);
/* Wait for all work of the handle. */
int ppk_mhd_synchronize(ppk_mhd_t *h);

/*
 * Per-stage GPU timers, the counterpart of the reference's phase timers (TIMER_BOUNDARIES,
 * TIMER_DT, TIMER_NUM_SCHEME; src/shared/SolverBase.h:35-42, 150-157): when a stage's bit is
 * set in stage_mask, every launch of that stage is bracketed by a pair of hipEvents recorded
 * on the handle's stream.  Enabling (or disabling) clears the accumulated times.
 */
int ppk_mhd_timers_enable(ppk_mhd_t *h, uint32_t stage_mask);
/* Sum of the recorded durations (ms) and the number of launches of one stage (synchronizes). */
int ppk_mhd_timers_read(ppk_mhd_t *h, int stage, double *total_ms, int64_t *launches);
int ppk_mhd_timer_count(void);
const char *ppk_mhd_timer_name(int stage);

/* Diagnostics (not on the reference path): max over interior cells of |div B| computed
 * from the face-centred field, and the interior sums of the 8 conservative variables.
 * Synchronizes. */
int ppk_mhd_divb_max(ppk_mhd_t *h, double *divb_max);
int ppk_mhd_conserved_sums(ppk_mhd_t *h, double sums[8]);
/* Device arithmetic self-check (no counterpart in the reference): for n host values x[i],
 * out[3i], out[3i+1], out[3i+2] receive 1/x, sqrt(x) and 1/sqrt(x) as the default FP64 build
 * computes them on the current device (v_rcp_f64 / v_rsq_f64 refined in FP64 arithmetic), so
 * that a caller can compare them with the IEEE operations.  Allocates device memory and
 * synchronizes. */
int ppk_mhd_math_selfcheck(const double *x, int64_t n, double *out);

/* ---- output ---------------------------------------------------------------------------- */
/* Write a state (host array, reference layout) as VTK ImageData <dir>/<prefix>_<istep>.vti
 * (istep zero-padded to 7 digits), byte for byte as the reference does: CellData arrays rho,
 * energy, rho_vx, rho_vy, rho_vz, bx, by, bz over the interior cells (the first nbvar of them:
 * 8 for MHD parameters, ghost_width 3; 4 / 5 for hydro parameters, ghost_width 2), appended raw
 * binary (ascii = 0) or ascii text (ascii = 1, the reference's [output] outputVtkAscii).
 * Replaces: save_VTK_2D / save_VTK_3D (src/utils/io/IO_VTK.cpp:23-410) with the variable
 * names of src/shared/SolverBase.cpp:48-55. */
int ppk_mhd_write_vtk(const ppk_mhd_params *params, const double *U, const char *dir,
                      const char *prefix, int istep, int ascii);
/* The same for the handle's current state (copies it to the host; synchronizes).
 * Replaces: SolverMHDMuscl::save_solution_impl (src/muscl/SolverMHDMuscl.h:894-906). */
int ppk_mhd_save_vtk(ppk_mhd_t *h, const char *dir, const char *prefix, int istep, int ascii);

/* Restart files: the state with its ghost cells, "time step" (the caller's output counter),
 * the iteration and "total time", in one raw little-endian file (layout in io_restart.cpp).
 * Replaces: the HDF5 restart output and reader (src/utils/io/IO_HDF5.h:430-484, attributes
 * "time step" and "total time"; SolverMHDMuscl::init_restart, src/muscl/SolverMHDMuscl.h:
 * 613-643).  A file whose grid differs from params is refused (PPK_ERR_INVALID_ARGUMENT). */
int ppk_mhd_write_restart(const ppk_mhd_params *params, const double *U, double total_time,
                          int64_t time_step, int64_t iteration, const char *path);
int ppk_mhd_read_restart(const ppk_mhd_params *params, const char *path, double *U, size_t count,
                         double *total_time, int64_t *time_step, int64_t *iteration);
/* The handle's current state and counters to a restart file, and back (synchronize).  Loading
 * restores t and the iteration; continuing gives the same results, bit for bit, as the run
 * that was never stopped.  (The reference restarts its iteration count at 0; a caller that
 * wants that calls ppk_mhd_set_time(h, t, 0) after loading.) */
int ppk_mhd_save_state(ppk_mhd_t *h, const char *path, int64_t time_step);
int ppk_mhd_load_state(ppk_mhd_t *h, const char *path, int64_t *time_step);

/* ---- batched ensembles ----------------------------------------------------------------------
 * Many independent simulations of one dimension and precision (any sizes, gamma, cfl,
 * boundaries, tEnd, n_step_max) advanced together: every stage of a step is one launch over
 * all members, and steps after the first replay one captured hipGraph.  Each member follows
 * its own time loop exactly as one ppk_mhd_t would with dt_mode PPK_DT_CFL (its own CFL dt,
 * clamp to tEnd, finished() test); the per-cell code is the fused path's, so its state after
 * each step equals that handle's bit for bit in PPK_FP64_STRICT (to roundoff in the other
 * precisions, where the compiler's FMA contractions may differ between the two kernels).
 * Replaces: a loop over the reference's driver, one SolverMHDMuscl per member
 * (src/main.cpp:130-160 with src/shared/SolverBase.cpp:161-219), for PPK_PATH_FUSED. */
typedef struct ppk_mhd_batch ppk_mhd_batch;
typedef ppk_mhd_batch ppk_mhd_batch_t;
/* params: nmembers parameter blocks (kernel_path is ignored: the batch always runs fused). */
int ppk_mhd_batch_create(const ppk_mhd_params *params, int nmembers, ppk_mhd_batch_t **handle);
void ppk_mhd_batch_destroy(ppk_mhd_batch_t *h);
const char *ppk_mhd_batch_last_error(const ppk_mhd_batch_t *h);
int ppk_mhd_batch_size(const ppk_mhd_batch_t *h);
/* problems: nmembers problem descriptions; t = 0, iteration = 0 for every member.
 * Replaces: SolverMHDMuscl::init (src/muscl/SolverMHDMuscl.h:868-889), per member. */
int ppk_mhd_batch_init_problem(ppk_mhd_batch_t *h, const ppk_mhd_problem *problems);
/* Member state in the reference layout (count = 8*isize*jsize*ksize of that member). */
int ppk_mhd_batch_upload(ppk_mhd_batch_t *h, int member, const double *U, size_t count);
int ppk_mhd_batch_download(ppk_mhd_batch_t *h, int member, double *U, size_t count);
/* nsteps steps of every member that has not finished; synchronizes at the end.
 * Replaces: SolverBase::next_iteration (src/shared/SolverBase.cpp:161-219), per member. */
int ppk_mhd_batch_run(ppk_mhd_batch_t *h, int nsteps);
int ppk_mhd_batch_get_time(ppk_mhd_batch_t *h, int member, double *t, double *dt,
                           int64_t *iteration);
int ppk_mhd_batch_synchronize(ppk_mhd_batch_t *h);
/* Block organisation of the batched kernels: tile = half-step tile shape (2D: 0 = 32x8,
 * 1 = 64x4, 2 = 16x16; 3D: 0 = 16x4x4, 1 = 8x8x4, 2 = 32x4x2), brick = cell order of the
 * update and face/edge kernels (0 = flattened rows in 4x4 patches; 2D: 1 = 64x4 bricks; 3D:
 * 1 = 64x2x2, 2 = 16x4x4, 3 = 32x4x2).  Results do not depend on the choice.  No counterpart
 * in the reference (Kokkos picks its own team shapes). */
int ppk_mhd_batch_set_tiles(ppk_mhd_batch_t *h, int tile, int brick);
int ppk_mhd_batch_get_tiles(const ppk_mhd_batch_t *h, int *tile, int *brick);
/* Autotuner: times trial_steps replays of the step for every tile x brick candidate on the
 * batch's own members, restores every member's state and time afterwards, and keeps the
 * fastest (returned with its ms per step).  No counterpart in the reference. */
int ppk_mhd_batch_autotune(ppk_mhd_batch_t *h, int trial_steps, int *tile, int *brick,
                           double *ms_per_step);

/* ---- surrogate form of the face solves (learned HLLD flux) -----------------------------------
 * No counterpart in the reference: ppkMHD solves every face with HLLD
 * (src/shared/RiemannSolvers_MHD.h:133-367).  A network 16 -> 128 -> 128 -> 8 (ReLU, bf16
 * weights and activations, FP32 accumulation on the matrix cores) learns the difference
 * between the HLLD and LLF fluxes of a face in dimensionless form; a face whose density and
 * pressure jumps are below the gate thresholds takes LLF + the learned correction if that flux
 * keeps the neighbouring cell updates positive, and exact HLLD otherwise.  Training samples
 * are dumped from the solver's own faces; training runs on the GPU with the library's own
 * forward, backward and Adam kernels.  All calls below are synchronous. */
typedef struct ppk_mlp ppk_mlp;
typedef ppk_mlp ppk_mlp_t;
typedef struct ppk_faceset ppk_faceset;
typedef ppk_faceset ppk_faceset_t;
/* FP32 parameter count: W1[128][16] b1[128] W2[128][128] b2[128] W3[8][128] b3[8], then the
 * standardisation mu_in[16] isig_in[16] mu_out[8] sig_out[8] (19768). */
int ppk_mlp_param_count(void);
/* He-uniform initial weights from `seed`, zero biases, identity standardisation. */
int ppk_mlp_create(uint64_t seed, ppk_mlp_t **net);
void ppk_mlp_destroy(ppk_mlp_t *net);
const char *ppk_mlp_last_error(const ppk_mlp_t *net);
int ppk_mlp_get_params(ppk_mlp_t *net, float *params, size_t count);
int ppk_mlp_set_params(ppk_mlp_t *net, const float *params, size_t count);
/* Training samples on the device: 16 features and 8 targets per face. */
int ppk_faceset_create(size_t capacity, ppk_faceset_t **set);
void ppk_faceset_destroy(ppk_faceset_t *set);
const char *ppk_faceset_last_error(const ppk_faceset_t *set);
int ppk_faceset_size(const ppk_faceset_t *set, size_t *n);
int ppk_faceset_clear(ppk_faceset_t *set);
int ppk_faceset_append(ppk_faceset_t *set, const float *X, const float *T, size_t n);
int ppk_faceset_read(ppk_faceset_t *set, size_t first, size_t n, float *X, float *T);
/* Append the samples of every stride-th face (from `phase`) that the solver's next CFL step
 * will solve: ghost fill, CFL dt, half step, then per face the features and the target
 * (F_HLLD - F_LLF) / scale.  Needs the fused path. */
int ppk_mhd_dump_faces(ppk_mhd_t *h, ppk_faceset_t *set, int stride, int phase, size_t *appended);
/* Feature and target means and standard deviations of a face set into the network. */
int ppk_mlp_standardize(ppk_mlp_t *net, const ppk_faceset_t *set);
/* `steps` Adam steps (lr, betas 0.9 / 0.999) on minibatches of `batch` samples (a multiple of
 * 256) drawn uniformly from the set with `seed`; the loss is the mean squared error of the
 * standardised outputs. */
int ppk_mlp_train(ppk_mlp_t *net, const ppk_faceset_t *set, int steps, int batch, float lr,
                  uint64_t seed);
/* Gradient of that loss on n given samples (n a multiple of 256) w.r.t. the 19720 trained
 * parameters, without updating them. */
int ppk_mlp_gradient(ppk_mlp_t *net, const float *X, const float *T, size_t n, float *grad,
                     size_t count);
/* Standardised outputs Y[n][8] of the network for features X[n][16]. */
int ppk_mlp_infer(ppk_mlp_t *net, const float *X, size_t n, float *Y);
/* Mean squared error per standardised output over a face set (8 doubles). */
int ppk_mlp_evaluate(ppk_mlp_t *net, const ppk_faceset_t *set, double *mse);
/* Use the network for the face solves of this solver (net = NULL: plain HLLD again).  The
 * network must outlive its use; gate thresholds are relative density / pressure jumps. */
int ppk_mhd_set_surrogate(ppk_mhd_t *h, const ppk_mlp_t *net, double gate_rho, double gate_p);
/* Faces solved by the network and by HLLD since the last reset. */
int ppk_mhd_surrogate_counts(ppk_mhd_t *h, int64_t *by_network, int64_t *by_hlld, int reset);

/* ---- Euler (hydro) MUSCL-Hancock solvers ------------------------------------------------------
 * The reference's Hydro_Muscl_2D / Hydro_Muscl_3D (src/muscl/SolverHydroMuscl.{h,cpp},
 * implementation version 0, no gravity): nbvar 4 in 2D (ID, IP, IU, IV) and 5 in 3D (+ IW),
 * ghost_width 2, same array layout as above with nbvar variables.  Parameters come in the
 * ppk_mhd_params struct with ghost_width 2, riemann approx, llf, hllc or (2D only) hll,
 * slope_type 0, 1 or 2, precision PPK_FP64 or PPK_FP64_STRICT, kernel_path staged or fused;
 * the initial condition in the ppk_mhd_problem struct, kind implode or blast.  The time-loop semantics (CFL dt clamped to tEnd,
 * U / U2 ping-pong, finished()) are those of the MHD handle.  PPK_FP64_STRICT results equal the
 * reference's double arithmetic bit for bit. */
typedef struct ppk_hydro_solver ppk_hydro_t;
/* Replaces: SolverHydroMuscl<dim>::SolverHydroMuscl allocation part
 * (src/muscl/SolverHydroMuscl.h:177-285), keys Hydro_Muscl_2D / Hydro_Muscl_3D
 * (src/shared/SolverFactory.cpp:26-27). */
int ppk_hydro_create(const ppk_mhd_params *params, ppk_hydro_t **handle);
void ppk_hydro_destroy(ppk_hydro_t *handle);
const char *ppk_hydro_last_error(const ppk_hydro_t *handle);
/* nbvar of the handle (4 or 5); 0 for NULL */
int ppk_hydro_nbvar(const ppk_hydro_t *handle);
/* Initial condition, ghost fill, U2 = U and the first dt, as the constructor does
 * (src/muscl/SolverHydroMuscl.h:286-300; implode HydroInitFunctors2D.h:35-125 /
 * HydroInitFunctors3D.h:66-165, blast HydroInitFunctors2D.h:129-212 / 3D.h:170-262). */
int ppk_hydro_init_problem(ppk_hydro_t *h, const ppk_mhd_problem *problem);
/* The initial condition alone, on the host (count = nbvar*isize*jsize*ksize). */
int ppk_hydro_initial_state(const ppk_mhd_params *params, const ppk_mhd_problem *problem,
                            double *U, size_t count);
int ppk_hydro_upload(ppk_hydro_t *h, const double *U, size_t count);
int ppk_hydro_download(ppk_hydro_t *h, double *U, size_t count);
int ppk_hydro_device_state(ppk_hydro_t *h, double **dU);
int ppk_hydro_stream(ppk_hydro_t *h, void **stream);
/* Replaces: SolverHydroMuscl::make_boundaries (src/muscl/SolverHydroMuscl.cpp:22-64). */
int ppk_hydro_make_boundaries(ppk_hydro_t *h);
/* Replaces: SolverHydroMuscl::compute_dt_local (src/muscl/SolverHydroMuscl.h:650-691). */
int ppk_hydro_compute_dt(ppk_hydro_t *h, double *dt);
/* One step; dt_mode as ppk_mhd_step.
 * Replaces: SolverHydroMuscl::next_iteration_impl -> godunov_unsplit_impl
 * (src/muscl/SolverHydroMuscl.h:697-760, SolverHydroMuscl.cpp:257-405). */
int ppk_hydro_step(ppk_hydro_t *h, double dt, int dt_mode);
/* nsteps steps, each a no-op once finished() holds (as ppk_mhd_run). */
int ppk_hydro_run(ppk_hydro_t *h, int nsteps, double dt, int dt_mode);
int ppk_hydro_get_time(ppk_hydro_t *h, double *t, double *dt, int64_t *iteration);
int ppk_hydro_set_time(ppk_hydro_t *h, double t, int64_t iteration);
/* Replaces: SolverBase::finished (src/shared/SolverBase.cpp:195-201). */
int ppk_hydro_finished(ppk_hydro_t *h, int *finished);
int ppk_hydro_synchronize(ppk_hydro_t *h);
/* The handle's current state as a VTK file (ppk_mhd_write_vtk with nbvar 4 / 5).
 * Replaces: SolverHydroMuscl::save_solution_impl (src/muscl/SolverHydroMuscl.h:809-820). */
int ppk_hydro_save_vtk(ppk_hydro_t *h, const char *dir, const char *prefix, int istep,
                       int ascii);
/* Per-stage GPU timers (hipEvents on the handle's stream), as ppk_mhd_timers_*; stages:
 * 0 boundaries, 1 primitives_dt, 2 fluxes, 3 update (staged form), 4 fused, 5 control.
 * Replaces: the per-stage timers of SolverBase (src/shared/SolverBase.h:35-42, 150-157). */
int ppk_hydro_timers_enable(ppk_hydro_t *h, uint32_t stage_mask);
int ppk_hydro_timers_read(ppk_hydro_t *h, int stage, double *total_ms, int64_t *launches);
int ppk_hydro_timer_count(void);
const char *ppk_hydro_timer_name(int stage);

#ifdef __cplusplus
}
#endif

#endif /* PPKMHD_AMD_H_ */
