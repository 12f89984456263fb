// On-device initial conditions: the per-cell functions of ics_body.hpp, one thread per cell,
// so no host array and no PCIe transfer.  Same operations and order as the host path; the
// transcendental functions of a cell (sin, cos, exp) come from the device math library.
#include <hip/hip_runtime.h>

#include "ics_body.hpp"

namespace ppk {
namespace icd {

constexpr int kBlock = 256;

__global__ void k_zero(double *U, long long n) {
  const long long t = (long long)blockIdx.x * kBlock + threadIdx.x;
  if (t < n) U[t] = 0.0;
}

// a box [lo, hi) of cells per direction, one thread per cell
struct Range {
  int i0, j0, k0, ni, nj, nk;
  __device__ bool cell(long long t, int &i, int &j, int &k) const {
    if (t >= (long long)ni * nj * nk) return false;
    i = i0 + (int)(t % ni);
    j = j0 + (int)((t / ni) % nj);
    k = k0 + (int)(t / ((long long)ni * nj));
    return true;
  }
  long long count() const { return (long long)ni * nj * nk; }
};

__global__ void k_cell(ppk_mhd_params p, ppk_mhd_problem pr, icb::IcConsts K, icb::IcBox b,
                       Range r, double *U) {
  int i, j, k;
  if (r.cell((long long)blockIdx.x * kBlock + threadIdx.x, i, j, k))
    icb::ic_cell(p, pr, K, b, i, j, k, U);
}
__global__ void k_ot_energy(ppk_mhd_params p, icb::IcConsts K, icb::IcBox b, Range r,
                            double *U) {
  int i, j, k;
  if (r.cell((long long)blockIdx.x * kBlock + threadIdx.x, i, j, k))
    icb::ic_ot_energy(p, K, b, i, j, k, U);
}
__global__ void k_wave_potential(ppk_mhd_params p, icb::IcConsts K, icb::IcBox b, Range r,
                                 double *A) {
  int i, j, k;
  if (r.cell((long long)blockIdx.x * kBlock + threadIdx.x, i, j, k))
    icb::ic_wave_potential(p, K, b, i, j, k, A);
}
__global__ void k_wave_field(ppk_mhd_params p, icb::IcBox b, Range r, const double *A,
                             double *U) {
  int i, j, k;
  if (r.cell((long long)blockIdx.x * kBlock + threadIdx.x, i, j, k))
    icb::ic_wave_field(p, b, i, j, k, A, U);
}
__global__ void k_wave_state(ppk_mhd_params p, icb::IcConsts K, icb::IcBox b, Range r,
                             double *U) {
  int i, j, k;
  if (r.cell((long long)blockIdx.x * kBlock + threadIdx.x, i, j, k))
    icb::ic_wave_state(p, K, b, i, j, k, U);
}

__global__ void k_loop_potential(ppk_mhd_params p, ppk_mhd_problem pr, icb::IcBox b, Range r,
                                 double *A) {
  int i, j, k;
  if (r.cell((long long)blockIdx.x * kBlock + threadIdx.x, i, j, k))
    icb::ic_loop_potential(p, pr, b, i, j, k, A);
}
__global__ void k_loop_state(ppk_mhd_params p, ppk_mhd_problem pr, icb::IcBox b, Range r,
                             const double *A, double *U) {
  int i, j, k;
  if (r.cell((long long)blockIdx.x * kBlock + threadIdx.x, i, j, k))
    icb::ic_loop_state(p, pr, b, i, j, k, A, U);
}
__global__ void k_loop_energy(ppk_mhd_params p, icb::IcBox b, Range r, double *U) {
  int i, j, k;
  if (r.cell((long long)blockIdx.x * kBlock + threadIdx.x, i, j, k))
    icb::ic_loop_energy(p, b, i, j, k, U);
}

inline dim3 grid(long long n) { return dim3((unsigned)((n + kBlock - 1) / kBlock)); }

}  // namespace icd

// U: 8 x cells (the state), A: 3 x cells of scratch (linear wave and field loop only; may
// alias nothing else)
hipError_t init_problem_device(const ppk_mhd_params &p, const ppk_mhd_problem &pr, double *U,
                               double *A, hipStream_t s) {
  using namespace icd;
  const icb::IcBox b = icb::box_of(p);
  const icb::IcConsts K = ic_consts(p, pr);
  const dim3 blk(kBlock);
  hipLaunchKernelGGL(k_zero, grid(8 * b.cells()), blk, 0, s, U, 8 * b.cells());
  const Range all{0, 0, 0, b.is, b.js, b.ks};
  if (pr.kind == PPK_PROBLEM_WAVE) {
    if (p.dim == 2) return hipGetLastError();
    hipLaunchKernelGGL(k_zero, grid(3 * b.cells()), blk, 0, s, A, 3 * b.cells());
    hipLaunchKernelGGL(k_wave_potential, grid(all.count()), blk, 0, s, p, K, b, all, A);
    const Range f{b.gw - 1, b.gw - 1, b.gw - 1, b.is - 2 * b.gw + 2, b.js - 2 * b.gw + 2,
                  b.ks - 2 * b.gw + 2};
    hipLaunchKernelGGL(k_wave_field, grid(f.count()), blk, 0, s, p, b, f, (const double *)A, U);
    const Range in{b.gw, b.gw, b.gw, b.is - 2 * b.gw, b.js - 2 * b.gw, b.ks - 2 * b.gw};
    hipLaunchKernelGGL(k_wave_state, grid(in.count()), blk, 0, s, p, K, b, in, U);
    return hipGetLastError();
  }
  if (pr.kind == PPK_PROBLEM_FIELD_LOOP) {
    hipLaunchKernelGGL(k_loop_potential, grid(all.count()), blk, 0, s, p, pr, b, all, A);
    const Range in{b.gw, b.gw, p.dim == 3 ? b.gw : 0, b.is - 2 * b.gw, b.js - 2 * b.gw,
                   p.dim == 3 ? b.ks - 2 * b.gw : 1};
    hipLaunchKernelGGL(k_loop_state, grid(in.count()), blk, 0, s, p, pr, b, in, (const double *)A, U);
    hipLaunchKernelGGL(k_loop_energy, grid(in.count()), blk, 0, s, p, b, in, U);
    return hipGetLastError();
  }
  hipLaunchKernelGGL(k_cell, grid(all.count()), blk, 0, s, p, pr, K, b, all, U);
  if (pr.kind == PPK_PROBLEM_ORSZAG_TANG) {
    const Range e{0, 0, 0, b.is - 1, b.js - 1, b.ks};
    hipLaunchKernelGGL(k_ot_energy, grid(e.count()), blk, 0, s, p, K, b, e, U);
  }
  return hipGetLastError();
}

}  // namespace ppk
