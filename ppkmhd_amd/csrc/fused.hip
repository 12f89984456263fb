// Fused form of the MHD step: three stages per time step in place of the reference's seven.
//
//   k_half3d / k_half2d        LDS-tiled: primitives, edge electric field, transverse
//                              face-field slopes and the Hancock half step of a tile of cells;
//                              one compact record per cell (kRec3 = 38 / kRec2 = 26 doubles)
//                              instead of the 18 (3D) / 8 (2D) full face and edge states
//   k_riemann_mix{3,2}d         per cell of the sub-box, every face and edge solve of the step in
//                              one launch (one solve per thread keeps register use low): the
//                              Riemann problem of the cell's lower face in each direction, the 2D
//                              HLLD EMF of its lower edges, the states rebuilt from the records
//                              of the cell and its lower neighbours.  The right-side face values
//                              of a cell (AR, dARy, ...) are bit for bit the left-side values of
//                              its +1 neighbour, so records hold only the left ones.  With the
//                              surrogate, k_edges3d / k_edge2d solve the edges after the
//                              learned faces.
//   k_update                   flux update + constrained transport, ghost copy, and the CFL
//                              maximum of the new state (per-block partials) that the next
//                              step's dt is taken from
//
// The arithmetic is the staged kernels' (same operations in the same order, references in
// step_staged.hip), so the PPK_NS=strict build is bit-identical to them and to the reference's
// double arithmetic.  The per-thread bodies are host+device functions: a host build runs them
// block by block, phase by phase, to check indexing against the oracle without a GPU.
#include <hip/hip_runtime.h>

#include "fused_body.hpp"

namespace ppk {
namespace PPK_NS {
namespace fusedk {

// ---- kernels --------------------------------------------------------------------------------
constexpr int kRiemannBlock = 256;

__device__ __forceinline__ bool active(const double *ctrl) { return ctrl[CTRL_ACTIVE] != 0.0; }

template <class T, class R>
__global__ void __launch_bounds__(T::NT) k_half3d(const double *__restrict__ U,
                                                  R *__restrict__ rec, DevParams p,
                                                  const double *ctrl, int ntx, int nty, int ntz) {
  if (!active(ctrl)) return;
  __shared__ R lds[T::LDS];
  R *sF = lds, *sQ = lds + 3 * T::NF, *sE = sQ + 8 * T::NQ;
  // tiles: x fastest, (y, z) tile rows in patches, each XCD a contiguous range
  const int L = xcd_remap(blockIdx.x, gridDim.x);
  int ty, tz;
  patch_row(L / ntx, kPatch, kPatch, (nty + kPatch - 1) / kPatch, ty, tz);
  if (ty >= nty || tz >= ntz) return;  // padding of the last patches
  const int b0x = p.gw - 1 + (L % ntx) * T::TX;
  const int b0y = p.gw - 1 + ty * T::TY;
  const int b0z = p.gw - 1 + tz * T::TZ;
  const double dt = ctrl[CTRL_DT];
  const int tid = threadIdx.x;
  half3_load<T, R>(U, p, b0x, b0y, b0z, tid, sF, sQ);
  __syncthreads();
  half3_prims<T, R>(p, b0x, b0y, b0z, tid, sF, sQ);
  __syncthreads();
  half3_elec<T, R>(p, b0x, b0y, b0z, tid, sF, sQ, sE);
  __syncthreads();
  half3_record<T, R>(p, b0x, b0y, b0z, tid, sF, sQ, sE, dt, rec);
}

template <class R>
__global__ void __launch_bounds__(Tile2::NT) k_half2d(const double *__restrict__ U,
                                                      R *__restrict__ rec, DevParams p,
                                                      const double *ctrl, int ntx) {
  if (!active(ctrl)) return;
  __shared__ R lds[Tile2::LDS];
  R *sF = lds, *sQ = lds + 2 * Tile2::NF, *sE = sQ + 8 * Tile2::NQ;
  const int L = xcd_remap(blockIdx.x, gridDim.x);
  const int b0x = p.gw - 1 + (L % ntx) * Tile2::TX;
  const int b0y = p.gw - 1 + (L / ntx) * Tile2::TY;
  const double dt = ctrl[CTRL_DT];
  const int tid = threadIdx.x;
  half2_load<Tile2, R>(U, p, b0x, b0y, tid, sF, sQ);
  __syncthreads();
  half2_prims<Tile2, R>(p, b0x, b0y, tid, sF, sQ);
  __syncthreads();
  half2_elec<Tile2, R>(p, b0x, b0y, tid, sF, sQ, sE);
  __syncthreads();
  half2_record<Tile2, R>(p, b0x, b0y, tid, sF, sQ, sE, dt, rec);
}

// The three edge solves in one launch (the surrogate step, whose faces are solved apart):
// block B does edge B % 3 (z, y, x) of run B / 3, renumbered per XCD like k_riemann_mix3d so
// that the three edges of a run read their records from one XCD's L2.  4 waves per SIMD: the
// EMF solve fits 128 VGPRs without spilling.
template <class R>
__global__ void __launch_bounds__(kRiemannBlock) __attribute__((amdgpu_waves_per_eu(4)))
k_edges3d(const R *__restrict__ rec, double *__restrict__ emf, DevParams p, Brick bk,
          const double *ctrl) {
  if (!active(ctrl)) return;
  const int B = xcd_remap(blockIdx.x, gridDim.x);
  const long long c = sub_brick_cell(p, bk, B / 3, threadIdx.x);
  if (c < 0) return;
  switch (B % 3) {
    case 0: edge3<2>(rec, emf, p, c); break;
    case 1: edge3<1>(rec, emf, p, c); break;
    default: edge3<0>(rec, emf, p, c); break;
  }
}
// All solves of the step in one launch: block B does solve type B % 6 (faces x, y, z, edges
// z, y, x) for the cells of run B / 6.  Blocks are renumbered per XCD first, so the six solves
// of a run execute on one XCD close in time: each record is read from HBM about once and then
// from that XCD's L2, and every CU interleaves the memory-bound face waves with the VALU-bound
// edge waves.  At 3 waves per SIMD the EMF solve needs no spill.  RS: the face solver (enum
// ppk_riemann) as a compile-time constant, so the kernel holds only the solver it runs.
template <int RS, class R>
__global__ void __launch_bounds__(kRiemannBlock) __attribute__((amdgpu_waves_per_eu(3)))
k_riemann_mix3d(const R *__restrict__ rec, R *__restrict__ flux, double *__restrict__ emf,
                DevParams p, Brick bk, const double *ctrl) {
  if (!active(ctrl)) return;
  const int B = xcd_remap(blockIdx.x, gridDim.x);
  const long long c = sub_brick_cell(p, bk, B / 6, threadIdx.x);
  if (c < 0) return;
  switch (B % 6) {
    case 0: face3<0, RS>(rec, flux, p, c); break;
    case 1: face3<1, RS>(rec, flux, p, c); break;
    case 2: face3<2, RS>(rec, flux, p, c); break;
    case 3: edge3<2>(rec, emf, p, c); break;
    case 4: edge3<1>(rec, emf, p, c); break;
    default: edge3<0>(rec, emf, p, c); break;
  }
}

template <int RS, class R>
__global__ void __launch_bounds__(kRiemannBlock) __attribute__((amdgpu_waves_per_eu(4)))
k_riemann_mix2d(const R *__restrict__ rec, R *__restrict__ flux,
                double *__restrict__ emf, DevParams p, Brick bk, const double *ctrl) {
  if (!active(ctrl)) return;
  const int B = xcd_remap(blockIdx.x, gridDim.x);
  const long long c = sub_brick_cell(p, bk, B / 3, threadIdx.x);
  if (c < 0) return;
  switch (B % 3) {
    case 0: face2<0, RS>(rec, flux, p, c); break;
    case 1: face2<1, RS>(rec, flux, p, c); break;
    default: edge2(rec, emf, p, c); break;
  }
}

template <class R>
__global__ void __launch_bounds__(kRiemannBlock) __attribute__((amdgpu_waves_per_eu(4)))
k_edge2d(const R *__restrict__ rec, double *__restrict__ emf, DevParams p, Brick bk,
         const double *ctrl) {
  if (!active(ctrl)) return;
  const long long c = sub_brick_cell(p, bk, xcd_remap(blockIdx.x, gridDim.x), threadIdx.x);
  if (c >= 0) edge2(rec, emf, p, c);
}

// one block per brick of the padded box; the block's CFL partial goes to parts[blockIdx.x]
template <class R>
__global__ void __launch_bounds__(kCflBlock) k_update(const double *__restrict__ Uin,
                                                      double *__restrict__ Uout,
                                                      const R *__restrict__ flux,
                                                      const double *__restrict__ emf,
                                                      double *__restrict__ parts, DevParams p,
                                                      Brick bk, const double *ctrl, int fold) {
  if (!active(ctrl)) return;
  __shared__ double red[kCflBlock / 64];
  const double dt = ctrl[CTRL_DT];
  const double dtdx = dt / p.dx, dtdy = dt / p.dy, dtdz = p.dim == 3 ? dt / p.dz : 0.0;
  int i, j, k;
  double best = 0.0;
  if (brick_cell(bk, xcd_remap(blockIdx.x, gridDim.x), threadIdx.x, i, j, k))
    best = update_cell(Uin, Uout, flux, emf, p, cidx(p, i, j, k), dtdx, dtdy, dtdz, fold != 0);
  if (!fold) return;
  best = block_max(best, red);
  if (threadIdx.x == 0) parts[blockIdx.x] = best;
}

__global__ void __launch_bounds__(kCflBlock) k_cfl(const double *__restrict__ U,
                                                   double *__restrict__ parts, DevParams p,
                                                   Brick bk, const double *ctrl) {
  if (!active(ctrl)) return;
  __shared__ double red[kCflBlock / 64];
  int i, j, k;
  double best = 0.0;
  if (brick_cell(bk, xcd_remap(blockIdx.x, gridDim.x), threadIdx.x, i, j, k))
    best = cfl_cell(U, p, cidx(p, i, j, k));
  best = block_max(best, red);
  if (threadIdx.x == 0) parts[blockIdx.x] = best;
}

inline int ceil_div(long long a, long long b) { return (int)((a + b - 1) / b); }

// the interleaved face/edge launch with the solver as a compile-time constant
template <class R>
void launch_mix3d(const DevParams &p, const R *rec, R *flux, double *emf, const double *ctrl,
                  hipStream_t s) {
  const Brick bk = sub_bricks(p);
  const dim3 g((unsigned)(6 * bk.count())), b(kRiemannBlock);
  switch (p.riemann) {
    case PPK_RIEMANN_HLLD:
      hipLaunchKernelGGL((k_riemann_mix3d<PPK_RIEMANN_HLLD, R>), g, b, 0, s, rec, flux, emf, p,
                         bk, ctrl);
      break;
    case PPK_RIEMANN_HLL:
      hipLaunchKernelGGL((k_riemann_mix3d<PPK_RIEMANN_HLL, R>), g, b, 0, s, rec, flux, emf, p,
                         bk, ctrl);
      break;
    default:
      hipLaunchKernelGGL((k_riemann_mix3d<PPK_RIEMANN_LLF, R>), g, b, 0, s, rec, flux, emf, p,
                         bk, ctrl);
      break;
  }
}
template <class R>
void launch_mix2d(const DevParams &p, unsigned grid, const R *rec, R *flux, double *emf, Brick bk,
                  const double *ctrl, hipStream_t s) {
  const dim3 g(grid), b(kRiemannBlock);
  switch (p.riemann) {
    case PPK_RIEMANN_HLLD:
      hipLaunchKernelGGL((k_riemann_mix2d<PPK_RIEMANN_HLLD, R>), g, b, 0, s, rec, flux, emf, p,
                         bk, ctrl);
      break;
    case PPK_RIEMANN_HLL:
      hipLaunchKernelGGL((k_riemann_mix2d<PPK_RIEMANN_HLL, R>), g, b, 0, s, rec, flux, emf, p,
                         bk, ctrl);
      break;
    default:
      hipLaunchKernelGGL((k_riemann_mix2d<PPK_RIEMANN_LLF, R>), g, b, 0, s, rec, flux, emf, p,
                         bk, ctrl);
      break;
  }
}

// the half step over every record tile: x fastest, (y, z) tile rows in 4 x 4 patches
template <class R>
void launch_half3(const DevParams &p, const double *Uin, R *rec, const double *ctrl,
                  hipStream_t s) {
  using T = Tile3;
  const int ntx = ceil_div(p.nx + 2, T::TX), nty = ceil_div(p.ny + 2, T::TY);
  const int ntz = ceil_div(p.nz + 2, T::TZ);
  const int prow = ceil_div(nty, kPatch) * kPatch * kPatch;  // tile rows per z-patch
  hipLaunchKernelGGL((k_half3d<T, R>), dim3(ntx * prow * ceil_div(ntz, kPatch)), dim3(T::NT), 0,
                     s, Uin, rec, p, ctrl, ntx, nty, ntz);
}

// the step's three stages in arithmetic type R (double, or float for the FP32-bulk form);
// rec and flux hold R, emf holds double
template <class R>
hipError_t godunov_impl(const DevParams &p, const double *Uin, double *Uout, R *rec, R *flux,
                        double *emf, double *parts, const double *ctrl, int fold_cfl,
                        const SurrogateCfg *sur, hipStream_t s, StageTimer *tm) {
  const Brick bk = sub_bricks(p);
  const dim3 g((unsigned)bk.count()), b(kRiemannBlock);
  if (p.dim == 3) {
    PPK_TIMED(tm, ST_FUSED_HALF, launch_half3<R>(p, Uin, rec, ctrl, s));
    if (tm) tm->mark(ST_FUSED_RIEMANN, true);
    if (sur) {  // learned face fluxes, then the exact edges
      const hipError_t e = surrogate_faces(p, rec, flux, ctrl, *sur, s, tm);
      if (e != hipSuccess) return e;
      hipLaunchKernelGGL((k_edges3d<R>), dim3((unsigned)(3 * bk.count())), b, 0, s, rec, emf, p,
                         bk, ctrl);
    } else {
      launch_mix3d<R>(p, rec, flux, emf, ctrl, s);
    }
    if (tm) tm->mark(ST_FUSED_RIEMANN, false);
  } else {
    const int ntx = ceil_div(p.nx + 2, Tile2::TX), nty = ceil_div(p.ny + 2, Tile2::TY);
    PPK_TIMED(tm, ST_FUSED_HALF,
              hipLaunchKernelGGL((k_half2d<R>), dim3(ntx * nty), dim3(Tile2::NT), 0, s, Uin, rec,
                                 p, ctrl, ntx));
    if (tm) tm->mark(ST_FUSED_RIEMANN, true);
    if (sur) {
      const hipError_t e = surrogate_faces(p, rec, flux, ctrl, *sur, s, tm);
      if (e != hipSuccess) return e;
      hipLaunchKernelGGL((k_edge2d<R>), g, b, 0, s, rec, emf, p, bk, ctrl);
    } else {
      launch_mix2d<R>(p, (unsigned)(3 * bk.count()), rec, flux, emf, bk, ctrl, s);
    }
    if (tm) tm->mark(ST_FUSED_RIEMANN, false);
  }
  PPK_TIMED(tm, ST_FUSED_UPDATE,
            hipLaunchKernelGGL((k_update<R>), dim3(fused_cfl_parts(p)), dim3(kCflBlock), 0, s,
                               Uin, Uout, flux, emf, parts, p, padded_bricks(p), ctrl, fold_cfl));
  return hipGetLastError();
}

}  // namespace fusedk

// -------------------------------------------------------------------------------------------
// launchers
// -------------------------------------------------------------------------------------------
hipError_t fused_godunov(const DevParams &p, const double *Uin, double *Uout, FusedScratch &w,
                         double *parts, const double *ctrl, int fold_cfl, hipStream_t s,
                         StageTimer *tm) {
  using namespace fusedk;
  if (p.bulk32)
    return godunov_impl<float>(p, Uin, Uout, static_cast<float *>(w.rec),
                               static_cast<float *>(w.flux), static_cast<double *>(w.emf), parts,
                               ctrl, fold_cfl, w.sur, s, tm);
  return godunov_impl<double>(p, Uin, Uout, static_cast<double *>(w.rec),
                              static_cast<double *>(w.flux), static_cast<double *>(w.emf), parts,
                              ctrl, fold_cfl, w.sur, s, tm);
}

template <class R>
hipError_t half_impl(const DevParams &p, const double *U, R *rec, const double *ctrl,
                     hipStream_t s) {
  using namespace fusedk;
  if (p.dim == 3) {
    launch_half3<R>(p, U, rec, ctrl, s);
  } else {
    const int ntx = ceil_div(p.nx + 2, Tile2::TX), nty = ceil_div(p.ny + 2, Tile2::TY);
    hipLaunchKernelGGL((k_half2d<R>), dim3(ntx * nty), dim3(Tile2::NT), 0, s, U, rec, p, ctrl,
                       ntx);
  }
  return hipGetLastError();
}

// the half-step records of U alone (the surrogate's training-sample dump reads them)
hipError_t fused_half(const DevParams &p, const double *U, void *rec, const double *ctrl,
                      hipStream_t s) {
  if (p.bulk32) return half_impl<float>(p, U, static_cast<float *>(rec), ctrl, s);
  return half_impl<double>(p, U, static_cast<double *>(rec), ctrl, s);
}

hipError_t fused_cfl(const DevParams &p, const double *U, double *parts, const double *ctrl,
                     hipStream_t s, StageTimer *tm) {
  using namespace fusedk;
  PPK_TIMED(tm, ST_FUSED_CFL,
            hipLaunchKernelGGL(k_cfl, dim3(fused_cfl_parts(p)), dim3(kCflBlock), 0, s, U, parts, p,
                               padded_bricks(p), ctrl));
  return hipGetLastError();
}

}  // namespace PPK_NS
}  // namespace ppk
