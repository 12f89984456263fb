// Interface between the host driver (capi.cpp) and the kernel translation units.  Each kernel
// file is compiled once per arithmetic variant into its own namespace (ppk::fast,
// ppk::strict); the driver picks one table of entry points per solver handle.
#pragma once

#include <hip/hip_runtime.h>

#include "mhd_device.hpp"

namespace ppk {

// slots of the device-resident time-loop control block (doubles)
enum {
  CTRL_T = 0,        // current time
  CTRL_DT = 1,       // dt of the step in flight
  CTRL_INVDT = 2,    // CFL accumulator (bit pattern of a non-negative double, atomicMax)
  CTRL_ITER = 3,     // iteration counter (exact in a double up to 2^53)
  CTRL_TEND = 4,
  CTRL_NSTEPMAX = 5,
  CTRL_ACTIVE = 6,   // 0 once SolverBase::finished() held before a checked step
  CTRL_INVDT_NEXT = 7,  // CFL accumulator of the state a fused step wrote
  CTRL_SLOTS = 8
};

// kernel stages bracketed by the optional per-stage GPU timers (ppk_mhd_timers_*)
enum Stage {
  ST_BOUNDARIES = 0,
  ST_PRIM_DT,
  ST_ELEC,
  ST_MAGSLOPES,
  ST_TRACE,
  ST_FLUXES,
  ST_EMF,
  ST_UPDATE,
  ST_CONTROL,
  ST_FUSED_HALF,     // fused: primitives + E + slopes + half step (compact trace data)
  ST_FUSED_UPDATE,   // fused: Riemann solves + EMFs + update + CT (+ next CFL)
  ST_FUSED_CFL,      // fused: standalone CFL reduction
  ST_COUNT
};
// host-side hook the launchers call around each stage (implemented by the driver)
struct StageTimer {
  virtual void mark(int stage, bool start) = 0;
  virtual ~StageTimer() {}
};
#define PPK_TIMED(tm, stage, launch)   \
  do {                                 \
    if (tm) (tm)->mark((stage), true); \
    launch;                            \
    if (tm) (tm)->mark((stage), false); \
  } while (0)

// scratch of the stage-per-kernel form (all over the padded box, SoA)
struct StagedScratch {
  double *faces = nullptr;  // 2*dim face states x 8 vars
  double *edges = nullptr;  // 4*(3 in 3D, 1 in 2D) edge states x 8 vars
  double *flux = nullptr;   // dim fluxes x 8 vars
  double *emf = nullptr;    // 3 (3D) or 1 (2D) EMF components
  double *elec = nullptr;   // 3D: 3 E components
  double *dbf = nullptr;    // 3D: 6 transverse face-field slopes
};

#define PPK_DECLARE_KERNELS(NS)                                                                  \
  namespace NS {                                                                                 \
  hipError_t staged_boundaries(const DevParams &p, double *U, const double *ctrl, hipStream_t s, \
                               StageTimer *tm);                                                  \
  hipError_t staged_prim_dt(const DevParams &p, const double *U, double *Q, double *ctrl,        \
                            int do_dt, hipStream_t s, StageTimer *tm);                           \
  hipError_t staged_begin(double *ctrl, int check_finished, hipStream_t s);                      \
  hipError_t staged_set_dt(double *ctrl, double fixed_dt, int cfl_mode, double cfl,              \
                           hipStream_t s);                                                       \
  hipError_t staged_end(double *ctrl, hipStream_t s);                                            \
  hipError_t staged_godunov(const DevParams &p, const double *Uin, double *Uout,                 \
                            const double *Q, StagedScratch &w, const double *ctrl,               \
                            hipStream_t s, StageTimer *tm);                                      \
  }

PPK_DECLARE_KERNELS(fast)
PPK_DECLARE_KERNELS(strict)

// entry-point table of one arithmetic variant
struct KernelTable {
  decltype(&fast::staged_boundaries) boundaries;
  decltype(&fast::staged_prim_dt) prim_dt;
  decltype(&fast::staged_begin) begin;
  decltype(&fast::staged_set_dt) set_dt;
  decltype(&fast::staged_end) end;
  decltype(&fast::staged_godunov) godunov;
};

}  // namespace ppk
