// Fused (LDS-tiled) form of the MHD step.  Declarations shared with the host driver.
#pragma once

#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace ppk {

struct FusedScratch {
  double *half = nullptr;   // compact half-step data of the trace cells
  double *aux = nullptr;    // small per-launch buffers
  size_t half_bytes = 0;
};

// true when the fused kernels handle these parameters under the requested path
bool fused_supported(const DevParams &p, int precision, int kernel_path);
hipError_t fused_alloc(const DevParams &p, int precision, FusedScratch &w);
void fused_free(FusedScratch &w);
// whole step: ghost fill of Uin, dt (fixed or CFL, clamped to tEnd - t), update into Uout
hipError_t fused_step(const DevParams &p, double *Uin, double *Uout, FusedScratch &w,
                      double *ctrl, double fixed_dt, int cfl_mode, int precision, hipStream_t s,
                      StageTimer *tm);
// CFL accumulator of the current state into ctrl[CTRL_INVDT] (ghosts already filled)
hipError_t fused_cfl(const DevParams &p, const double *U, FusedScratch &w, double *ctrl,
                     hipStream_t s, StageTimer *tm);

}  // namespace ppk
