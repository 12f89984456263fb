// Interface between the hydro (Euler) driver (hydro_capi.cpp) and its kernels (hydro.hip),
// compiled once per arithmetic variant (ppk::fast, ppk::strict) like the MHD kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "hydro_body.hpp"
#include "kernels.hpp"

namespace ppk {

// Everything the hydro kernels read, passed by value.  nv = 4 in 2D (ID, IP, IU, IV), 5 in 3D.
struct HydroDev {
  int dim, nv, nx, ny, nz, gw;
  int isize, jsize, ksize;
  int bc[6];
  int strip_rows;  // fused kernels: output rows (2D) / planes (3D) one wave or block marches
  long long plane, vstride;
  double dx, dy, dz;
  double idx, idy, idz;  // 1 / dx ... (used only where the build shares reciprocals)
  hyd::HydroPar P;
  // gravity: grav != 0 adds the half-step predictor to every face state, the source term to
  // every updated cell and the gravity form of the CFL term, each with its own cell's field
  // gf (dim planes, vstride apart, padded-box layout)
  int grav;
  const double *gf;
  double cfl, dxmin, idxmin;  // the gravity CFL term works on the smallest cell size
};

// fused marching kernels: a wave covers 64 columns of x, of which the middle kHydroCols are its
// output (two columns of halo on each side)
constexpr int kHydroLanes = 64;
constexpr int kHydroCols = kHydroLanes - 4;
// 3D: a block of kHydroWaves3 waves; each wave holds 64 / kHydroRowW3 rows of kHydroRowW3
// columns, the rows of the block stacked along y.  The middle kHydroCols3 columns and
// kHydroRows3 rows are the block's output (two columns / rows of halo on each side): 32-wide
// rows give 28 x 12 output cells per block and plane where 64-wide ones gave 60 x 4.
constexpr int kHydroWaves3 = 8;
constexpr int kHydroRowW3 = 32;
constexpr int kHydroBlockRows3 = kHydroWaves3 * (kHydroLanes / kHydroRowW3);
constexpr int kHydroCols3 = kHydroRowW3 - 4;
constexpr int kHydroRows3 = kHydroBlockRows3 - 4;

#define PPK_DECLARE_HYDRO(NS)                                                                    \
  namespace NS {                                                                                 \
  /* ghost fill of U (x, then y, then z); with copy != nullptr every ghost value is also        \
     written there (the next state's ghosts = this state's, as the reference's deep copy) */    \
  hipError_t hydro_boundaries(const HydroDev &h, double *U, double *copy, const double *ctrl,    \
                              hipStream_t s);                                                    \
  /* primitives of every cell into Q (staged form) and per-block CFL maxima of the interior */   \
  hipError_t hydro_prim_dt(const HydroDev &h, const double *U, double *Q, double *parts,         \
                           const double *ctrl, int do_dt, hipStream_t s);                        \
  int hydro_prim_parts(const HydroDev &h);                                                       \
  /* staged form: fluxes of every left face of [gw, size-gw], then the update of Uout (which    \
     holds a copy of Uin) */                                                                     \
  hipError_t hydro_staged_fluxes(const HydroDev &h, const double *Q, double *F,                  \
                                 const double *ctrl, hipStream_t s);                             \
  hipError_t hydro_staged_update(const HydroDev &h, const double *F, double *Uout,               \
                                 const double *ctrl, hipStream_t s);                             \
  /* fused form: one marching kernel from Uin to the interior of Uout; with fold_cfl the        \
     per-wave / per-block CFL maxima of the new state go to parts */                             \
  hipError_t hydro_fused(const HydroDev &h, const double *Uin, double *Uout, double *parts,      \
                         const double *ctrl, int fold_cfl, hipStream_t s);                       \
  int hydro_fused_parts(const HydroDev &h);                                                      \
  }

PPK_DECLARE_HYDRO(fast)
PPK_DECLARE_HYDRO(strict)

struct HydroTable {
  decltype(&fast::hydro_boundaries) boundaries;
  decltype(&fast::hydro_prim_dt) prim_dt;
  decltype(&fast::hydro_prim_parts) prim_parts;
  decltype(&fast::hydro_staged_fluxes) staged_fluxes;
  decltype(&fast::hydro_staged_update) staged_update;
  decltype(&fast::hydro_fused) fused;
  decltype(&fast::hydro_fused_parts) fused_parts;
};

}  // namespace ppk
