// Single-pass 3D MHD step: one kernel reads the state once and writes the new state once
// (bodies and data flow: march_body.hpp).  The ghost cells of the new state are the
// reference's deep copy of the filled input ghosts (U2 = U before the update), copied by a
// small kernel over the same elements the ghost fill writes.
#include <hip/hip_runtime.h>

#include "march_body.hpp"

namespace ppk {
namespace PPK_NS {
namespace marchk {

static_assert(kMO == kMarchOut, "output tile edge of the march kernel");

// the block's program on the device: every thread runs each phase, then the block syncs
template <class T>
struct DevExec {
  T &t;
  __device__ int tid(const T &) const { return (int)threadIdx.x; }
  template <class F>
  __device__ void each(F &&f) {
    f(t);
    __syncthreads();
  }
  __device__ void cell_cfl(const T &, int, double) {}
};

// One block per CU, marching its share of the grid's tile-planes (tiles of 16 x 16 record
// columns); 4 waves, one block per CU (the LDS of one block is 153 KB), one wave per SIMD with
// up to 512 registers.
template <int RS>
__global__ void __launch_bounds__(kMN, 1) __attribute__((amdgpu_waves_per_eu(1, 1)))
k_march3d(const double *__restrict__ Uin, double *__restrict__ Uout, double *__restrict__ parts,
          DevParams p, MarchGeom g, const double *ctrl, int fold) {
  if (ctrl[CTRL_ACTIVE] == 0.0) return;
  __shared__ double lds[kMarchLds];
  __shared__ double red[kMN / 64];
  MarchThread<RS> t;
  DevExec<MarchThread<RS>> X{t};
  const int L = xcd_remap(blockIdx.x, gridDim.x);
  march_block<RS>(X, Uin, Uout, p, g, L, ctrl[CTRL_DT], fold != 0, lds);
  if (fold) {
    const double m = block_max(t.cfl, red);
    if (threadIdx.x == 0) parts[blockIdx.x] = m;
  }
}

// ghost cells of Uout = those of Uin, over the elements the ghost fill of direction dir writes
// (bnd_element's enumeration)
__global__ void __launch_bounds__(256) k_ghost_copy(const double *__restrict__ Uin,
                                                   double *__restrict__ Uout, DevParams p,
                                                   int dir, const double *ctrl) {
  if (ctrl[CTRL_ACTIVE] == 0.0) return;
  const long long n = bnd_elements(p, dir);
  for (long long t = (long long)blockIdx.x * 256 + threadIdx.x; t < n;
       t += (long long)gridDim.x * 256) {
    const int gw = p.gw;
    const int nn = dir == 0 ? p.nx : (dir == 1 ? p.ny : p.nz);
    const int s1 = dir == 0 ? p.jsize : p.isize;
    const int s2 = dir == 2 ? p.jsize : p.ksize;
    const int g = (int)(t % gw);
    long long r = t / gw;
    const int t1 = (int)(r % s1);
    r /= s1;
    const int t2 = (int)(r % s2);
    const int upper = (int)(r / s2);
    const int c = upper ? g + nn + gw : g;
    int i, j, k;
    if (dir == 0) { i = c; j = t1; k = t2; }
    else if (dir == 1) { i = t1; j = c; k = t2; }
    else { i = t1; j = t2; k = c; }
    const long long d = cidx(p, i, j, k);
#pragma unroll
    for (int v = 0; v < 8; ++v) Uout[d + v * p.vstride] = Uin[d + v * p.vstride];
  }
}

}  // namespace marchk

hipError_t march_godunov(const DevParams &p, const double *Uin, double *Uout, double *parts,
                         const double *ctrl, int fold_cfl, hipStream_t s, StageTimer *tm) {
  using namespace marchk;
  const MarchGeom g = march_geom(p, p.march);
  const dim3 grid((unsigned)g.blocks()), block(kMN);
  if (tm) tm->mark(ST_FUSED_MARCH, true);
  switch (p.riemann) {
    case PPK_RIEMANN_HLLD:
      hipLaunchKernelGGL((k_march3d<PPK_RIEMANN_HLLD>), grid, block, 0, s, Uin, Uout, parts, p,
                         g, ctrl, fold_cfl);
      break;
    case PPK_RIEMANN_HLL:
      hipLaunchKernelGGL((k_march3d<PPK_RIEMANN_HLL>), grid, block, 0, s, Uin, Uout, parts, p,
                         g, ctrl, fold_cfl);
      break;
    default:
      hipLaunchKernelGGL((k_march3d<PPK_RIEMANN_LLF>), grid, block, 0, s, Uin, Uout, parts, p,
                         g, ctrl, fold_cfl);
      break;
  }
  if (tm) tm->mark(ST_FUSED_MARCH, false);
  for (int dir = 0; dir < 3; ++dir) {
    const long long n = bnd_elements(p, dir);
    long long nb = (n + 255) / 256;
    if (nb > 4096) nb = 4096;
    hipLaunchKernelGGL(k_ghost_copy, dim3((unsigned)nb), dim3(256), 0, s, Uin, Uout, p, dir,
                       ctrl);
  }
  return hipGetLastError();
}

}  // namespace PPK_NS
}  // namespace ppk
