// Diagnostics outside the reference path: max |div B| from the face-centred field and the
// interior sums of the conservative variables (per-block partials, summed in a fixed order on
// the host so the result does not depend on scheduling).
#include <hip/hip_runtime.h>

#include "mhd_device.hpp"

namespace ppk {
namespace {

constexpr int kDiagThreads = 256;

__global__ void __launch_bounds__(kDiagThreads) k_divb(const double *__restrict__ U, DevParams p,
                                                       double *out) {
  const long long n = (long long)p.nx * p.ny * p.nz;
  double best = 0.0;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n;
       t += (long long)gridDim.x * blockDim.x) {
    int i = (int)(t % p.nx) + p.gw;
    long long r = t / p.nx;
    int j = (int)(r % p.ny) + p.gw;
    int k = (int)(r / p.ny) + (p.dim == 3 ? p.gw : 0);
    long long c = ((long long)k * p.jsize + j) * p.isize + i;
    double d = (U[c + 1 + IA * p.vstride] - U[c + IA * p.vstride]) / p.dx +
               (U[c + p.isize + IB * p.vstride] - U[c + IB * p.vstride]) / p.dy;
    if (p.dim == 3) d += (U[c + p.plane + IC * p.vstride] - U[c + IC * p.vstride]) / p.dz;
    best = fmax(best, fabs(d));
  }
  for (int off = 32; off > 0; off >>= 1) best = fmax(best, __shfl_xor(best, off, 64));
  if ((threadIdx.x & 63) == 0)
    atomicMax((unsigned long long *)out, (unsigned long long)__double_as_longlong(best));
}

__global__ void __launch_bounds__(kDiagThreads) k_sums(const double *__restrict__ U, DevParams p,
                                                       double *partials) {
  __shared__ double red[kDiagThreads / 64][8];
  const long long n = (long long)p.nx * p.ny * p.nz;
  double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n;
       t += (long long)gridDim.x * blockDim.x) {
    int i = (int)(t % p.nx) + p.gw;
    long long r = t / p.nx;
    int j = (int)(r % p.ny) + p.gw;
    int k = (int)(r / p.ny) + (p.dim == 3 ? p.gw : 0);
    long long c = ((long long)k * p.jsize + j) * p.isize + i;
#pragma unroll
    for (int v = 0; v < 8; ++v) acc[v] += U[c + v * p.vstride];
  }
#pragma unroll
  for (int v = 0; v < 8; ++v)
    for (int off = 32; off > 0; off >>= 1) acc[v] += __shfl_xor(acc[v], off, 64);
  const int wave = threadIdx.x / 64;
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int v = 0; v < 8; ++v) red[wave][v] = acc[v];
  __syncthreads();
  if (threadIdx.x < 8) {
    double s = 0.0;
    for (int w = 0; w < kDiagThreads / 64; ++w) s += red[w][threadIdx.x];
    partials[(long long)blockIdx.x * 8 + threadIdx.x] = s;
  }
}

// the default FP64 build's refined reciprocal, square root and inverse square root
__global__ void __launch_bounds__(kDiagThreads) k_math(const double *__restrict__ x, long long n,
                                                       double *__restrict__ out) {
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n;
       t += (long long)gridDim.x * blockDim.x) {
    const double v = x[t];
    out[3 * t] = fast_rcp(v);
    out[3 * t + 1] = fast_sqrt(v);
    out[3 * t + 2] = fast_rsq(v);
  }
}

__global__ void k_zero1(double *x) {
  if (threadIdx.x == 0) x[0] = 0.0;
}

}  // namespace

hipError_t diag_divb(const DevParams &p, const double *U, double *out, hipStream_t s) {
  hipLaunchKernelGGL(k_zero1, dim3(1), dim3(64), 0, s, out);
  hipLaunchKernelGGL(k_divb, dim3(512), dim3(kDiagThreads), 0, s, U, p, out);
  return hipGetLastError();
}

hipError_t diag_math(const double *x, long long n, double *out, hipStream_t s) {
  hipLaunchKernelGGL(k_math, dim3(1024), dim3(kDiagThreads), 0, s, x, n, out);
  return hipGetLastError();
}

hipError_t diag_sums(const DevParams &p, const double *U, double *partials, int nblocks,
                     hipStream_t s) {
  hipLaunchKernelGGL(k_sums, dim3(nblocks), dim3(kDiagThreads), 0, s, U, p, partials);
  return hipGetLastError();
}

}  // namespace ppk
