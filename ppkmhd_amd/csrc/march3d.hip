// Single-pass 3D MHD step: one kernel reads the state once and writes the interior of the new
// state once (bodies and data flow: march_body.hpp).  The ghost cells of the new state are the
// reference's deep copy of the filled input ghosts (U2 = U before the update): the step's ghost
// fill (staged_boundaries with the new state as its copy target) writes each ghost value to
// both states.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "march_body.hpp"

namespace ppk {
namespace PPK_NS {
namespace marchk {

static_assert(kMOX == kMarchOutX && kMOY == kMarchOutY, "output tile of the march kernel");

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

// One block per CU, marching its share of the grid's tile-planes (tiles of 18 x 14 record
// columns); 4 waves, one block per CU (the LDS of one block is 152 KB), one wave per SIMD with
// up to 512 registers.
// RS: Riemann solver; ST: 1 for minmod slopes with the factor fixed, 0 to read slope_type;
// ISO: 0 when the equation of state is not isothermal (cIso <= 0), so that the isothermal
// branches fold away (the kernel works on a copy of the parameters with cIso = 0, which takes
// the same branches as any cIso <= 0).
template <int RS, int ST, int ISO>
__global__ void __launch_bounds__(kMN, 1) __attribute__((amdgpu_waves_per_eu(1, 1)))
k_march3d(const double *__restrict__ Uin, double *__restrict__ Uout, double *__restrict__ parts,
          DevParams pk, MarchGeom g, const double *ctrl, int fold) {
  if (ctrl[CTRL_ACTIVE] == 0.0) return;
  DevParams p = pk;
  if (!ISO) p.cIso = 0.0;
  __shared__ double lds[kMarchLds];
  __shared__ double red[kMN / 64];
  MarchThread<RS, ST> t;
  DevExec<MarchThread<RS, ST>> X{t};
  const int L = xcd_remap(blockIdx.x, gridDim.x);
  march_block<RS, ST>(X, Uin, Uout, p, g, L, ctrl[CTRL_DT], fold != 0, lds);
  if (fold) {
    const double m = block_max(t.cfl, red);
    if (threadIdx.x == 0) parts[blockIdx.x] = m;
  }
}

}  // namespace marchk

hipError_t march_godunov(const DevParams &p, const double *Uin, double *Uout, double *parts,
                         const double *ctrl, int fold_cfl, hipStream_t s, StageTimer *tm) {
  using namespace marchk;
  const MarchGeom g = march_geom(p, p.march);
  const dim3 grid((unsigned)g.blocks()), block(kMN);
  if (tm) tm->mark(ST_FUSED_MARCH, true);
  auto launch = [&](auto kernel) {
    hipLaunchKernelGGL(kernel, grid, block, 0, s, Uin, Uout, parts, p, g, ctrl, fold_cfl);
  };
  // minmod slopes (slope_type 1) with the factor fixed at compile time (x * 1 is exact, so the
  // result is the same; two multiplies fewer per limited slope); other slope types read it.
  // Isothermal runs (cIso > 0) take the kernels that keep the isothermal branches.
  const int minmod = p.slope_type == 1.0 ? 1 : 0, iso = p.cIso > 0 ? 1 : 0;
  auto pick = [&](auto rs) {
    constexpr int RS = decltype(rs)::value;
    if (minmod)
      iso ? launch(k_march3d<RS, 1, 1>) : launch(k_march3d<RS, 1, 0>);
    else
      iso ? launch(k_march3d<RS, 0, 1>) : launch(k_march3d<RS, 0, 0>);
  };
  switch (p.riemann) {
    case PPK_RIEMANN_HLLD: pick(std::integral_constant<int, PPK_RIEMANN_HLLD>{}); break;
    case PPK_RIEMANN_HLL: pick(std::integral_constant<int, PPK_RIEMANN_HLL>{}); break;
    default: pick(std::integral_constant<int, PPK_RIEMANN_LLF>{}); break;
  }
  if (tm) tm->mark(ST_FUSED_MARCH, false);
  return hipGetLastError();
}

}  // namespace PPK_NS
}  // namespace ppk
