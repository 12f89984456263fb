// Stage-per-kernel form of the MHD step: one HIP kernel per stage of the reference's
// implementation 0 (ghost fill, primitives + CFL reduction, E field, magnetic slopes, trace,
// fluxes, EMFs, update + CT), SoA arrays over the ghost-padded box.  It is the structural
// twin of the reference and the yardstick the fused kernels are checked and timed against.
//
// This file is compiled twice: PPK_NS=fast (FMA contraction allowed) and PPK_NS=strict
// (-ffp-contract=off: bit-identical to the reference's double arithmetic).
#include <hip/hip_runtime.h>

#include "kernels.hpp"
#include "trace.hpp"

namespace ppk {
namespace PPK_NS {

namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ long long idx(const DevParams &p, int i, int j, int k) {
  return ((long long)k * p.jsize + j) * p.isize + i;
}

template <class R>
__device__ __forceinline__ void load8(const double *__restrict__ a, const DevParams &p, long long c,
                                      St<R> &s) {
#pragma unroll
  for (int v = 0; v < 8; ++v) s[v] = R(a[c + v * p.vstride]);
}
template <class R>
__device__ __forceinline__ void store8(double *__restrict__ a, const DevParams &p, long long c,
                                       const St<R> &s) {
#pragma unroll
  for (int v = 0; v < 8; ++v) a[c + v * p.vstride] = double(s[v]);
}

__device__ __forceinline__ bool active(const double *ctrl) { return ctrl[CTRL_ACTIVE] != 0.0; }

__device__ __forceinline__ void decompose(const DevParams &p, long long t, int &i, int &j, int &k) {
  i = (int)(t % p.isize);
  long long r = t / p.isize;
  j = (int)(r % p.jsize);
  k = (int)(r / p.jsize);
}

// ---- ghost fill of every face, edge and corner in one pass: ghost_shell_element
// (kernels.hpp) per ghost cell, the values of the reference's x, y, z passes
// (src/shared/SolverBase.cpp:339-537, src/shared/BoundariesFunctors.h:534-1057) ------------
__global__ void __launch_bounds__(kBlock) k_boundaries(double *__restrict__ U,
                                                       double *__restrict__ copy, DevParams p,
                                                       const double *ctrl) {
  if (!active(ctrl)) return;
  const long long total = ghost_shell_cells(p);
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < total;
       t += (long long)gridDim.x * blockDim.x)
    ghost_shell_element(U, p, t, copy);
}

// ---- primitives over [0, size-1) per direction + CFL max over the interior ----------------
// reference src/muscl/MHDRunFunctors3D.h:88-163 and :16-83 (reduction fused here); the max is
// left as one partial per block in parts[blockIdx.x] (grid = cfl_blocks)
__global__ void __launch_bounds__(kBlock) k_prim_dt(const double *__restrict__ U,
                                                    double *__restrict__ Q, DevParams p,
                                                    double *__restrict__ parts,
                                                    const double *ctrl, int do_dt) {
  if (!active(ctrl)) return;
  __shared__ double red[kBlock / 64];
  const long long total = (long long)p.isize * p.jsize * p.ksize;
  double best = 0.0;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < total;
       t += (long long)gridDim.x * blockDim.x) {
    int i, j, k;
    decompose(p, t, i, j, k);
    const bool in3 = p.dim == 3 ? (k < p.ksize - 1) : true;
    if (i < p.isize - 1 && j < p.jsize - 1 && in3) {
      St<double> u, q;
      load8(U, p, t, u);
      double mx = U[t + 1 + IA * p.vstride];
      double my = U[t + p.isize + IB * p.vstride];
      double mz = p.dim == 3 ? U[t + p.plane + IC * p.vstride] : 0.0;
      constoprim(u, mx, my, mz, q, p);
      store8(Q, p, t, q);
      const int g = p.gw;
      bool inner = i >= g && i < p.isize - g && j >= g && j < p.jsize - g &&
                   (p.dim == 2 || (k >= g && k < p.ksize - g));
      if (do_dt && inner) best = fmax(best, inv_dt_cell(q, p));
    }
  }
  if (!do_dt) return;
  best = block_max(best, red);
  if (threadIdx.x == 0) parts[blockIdx.x] = best;
}

// ---- E field at cell edges (3D), reference src/muscl/MHDRunFunctors3D.h:303-356 ----------
__global__ void __launch_bounds__(kBlock) k_elec3d(const double *__restrict__ U,
                                                   const double *__restrict__ Q,
                                                   double *__restrict__ E, DevParams p,
                                                   const double *ctrl) {
  if (!active(ctrl)) return;
  const long long total = (long long)p.isize * p.jsize * p.ksize;
  const long long vs = p.vstride, sx = 1, sy = p.isize, sz = p.plane;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < total;
       t += (long long)gridDim.x * blockDim.x) {
    int i, j, k;
    decompose(p, t, i, j, k);
    if (!(k > 0 && k < p.ksize - 1 && j > 0 && j < p.jsize - 1 && i > 0 && i < p.isize - 1))
      continue;
    auto q = [&](long long off, int v) { return Q[t + off + v * vs]; };
    auto u = [&](long long off, int v) { return U[t + off + v * vs]; };
    double v_, w_, u_, A, B, C;
    v_ = 0.25 * (q(-sy - sz, IV) + q(-sy, IV) + q(-sz, IV) + q(0, IV));
    w_ = 0.25 * (q(-sy - sz, IW) + q(-sy, IW) + q(-sz, IW) + q(0, IW));
    B = 0.5 * (u(-sz, IB) + u(0, IB));
    C = 0.5 * (u(-sy, IC) + u(0, IC));
    E[t + 0 * vs] = v_ * C - w_ * B;
    u_ = 0.25 * (q(-sx - sz, IU) + q(-sx, IU) + q(-sz, IU) + q(0, IU));
    w_ = 0.25 * (q(-sx - sz, IW) + q(-sx, IW) + q(-sz, IW) + q(0, IW));
    A = 0.5 * (u(-sz, IA) + u(0, IA));
    C = 0.5 * (u(-sx, IC) + u(0, IC));
    E[t + 1 * vs] = w_ * A - u_ * C;
    u_ = 0.25 * (q(-sx - sy, IU) + q(-sx, IU) + q(-sy, IU) + q(0, IU));
    v_ = 0.25 * (q(-sx - sy, IV) + q(-sx, IV) + q(-sy, IV) + q(0, IV));
    A = 0.5 * (u(-sy, IA) + u(0, IA));
    B = 0.5 * (u(-sx, IB) + u(0, IB));
    E[t + 2 * vs] = u_ * B - v_ * A;
  }
}

// ---- transverse face-field slopes (3D), reference src/muscl/MHDRunFunctors3D.h:472-532 ----
// layout: dB3[c*vs] with c = 0 dAy, 1 dAz, 2 dBx, 3 dBz, 4 dCx, 5 dCy
__global__ void __launch_bounds__(kBlock) k_magslopes3d(const double *__restrict__ U,
                                                        double *__restrict__ dbf, DevParams p,
                                                        const double *ctrl) {
  if (!active(ctrl)) return;
  const long long total = (long long)p.isize * p.jsize * p.ksize;
  const long long vs = p.vstride, sx = 1, sy = p.isize, sz = p.plane;
  const double xst = fmin(p.slope_type, 2.0);
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < total;
       t += (long long)gridDim.x * blockDim.x) {
    int i, j, k;
    decompose(p, t, i, j, k);
    if (!(k > 0 && k < p.ksize - 1 && j > 0 && j < p.jsize - 1 && i > 0 && i < p.isize - 1))
      continue;
    auto u = [&](long long off, int v) { return U[t + off + v * vs]; };
    dbf[t + 0 * vs] = limited_slope(xst, u(0, IA), u(sy, IA), u(-sy, IA));
    dbf[t + 1 * vs] = limited_slope(xst, u(0, IA), u(sz, IA), u(-sz, IA));
    dbf[t + 2 * vs] = limited_slope(xst, u(0, IB), u(sx, IB), u(-sx, IB));
    dbf[t + 3 * vs] = limited_slope(xst, u(0, IB), u(sz, IB), u(-sz, IB));
    dbf[t + 4 * vs] = limited_slope(xst, u(0, IC), u(sx, IC), u(-sx, IC));
    dbf[t + 5 * vs] = limited_slope(xst, u(0, IC), u(sy, IC), u(-sy, IC));
  }
}

// ---- 3D trace: face states (Qm/Qp x,y,z) and 12 edge states --------------------------------
// range [gw-2, size-gw+1)^3, reference src/muscl/MHDRunFunctors3D.h:669-843
__global__ void __launch_bounds__(kBlock) k_trace3d(const double *__restrict__ U,
                                                    const double *__restrict__ Q,
                                                    const double *__restrict__ E,
                                                    const double *__restrict__ dbf,
                                                    double *__restrict__ faces,
                                                    double *__restrict__ edges, DevParams p,
                                                    const double *ctrl) {
  if (!active(ctrl)) return;
  const double dt = ctrl[CTRL_DT];
  const double dtdx = dt / p.dx, dtdy = dt / p.dy, dtdz = dt / p.dz;
  const long long total = (long long)p.isize * p.jsize * p.ksize;
  const long long vs = p.vstride, sx = 1, sy = p.isize, sz = p.plane;
  const long long v8 = 8 * vs;
  const int lo = p.gw - 2;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < total;
       t += (long long)gridDim.x * blockDim.x) {
    int i, j, k;
    decompose(p, t, i, j, k);
    if (!(i >= lo && i < p.isize - p.gw + 1 && j >= lo && j < p.jsize - p.gw + 1 && k >= lo &&
          k < p.ksize - p.gw + 1))
      continue;
    St<double> q, qPX, qMX, qPY, qMY, qPZ, qMZ;
    load8(Q, p, t, q);
    load8(Q, p, t + sx, qPX);
    load8(Q, p, t - sx, qMX);
    load8(Q, p, t + sy, qPY);
    load8(Q, p, t - sy, qMY);
    load8(Q, p, t + sz, qPZ);
    load8(Q, p, t - sz, qMZ);
    double bfNb[6] = {U[t + IA * vs], U[t + sx + IA * vs], U[t + IB * vs],
                      U[t + sy + IB * vs], U[t + IC * vs], U[t + sz + IC * vs]};
    double d12[12] = {dbf[t + 0 * vs],      dbf[t + 1 * vs],      dbf[t + 2 * vs],
                      dbf[t + 3 * vs],      dbf[t + 4 * vs],      dbf[t + 5 * vs],
                      dbf[t + sx + 0 * vs], dbf[t + sx + 1 * vs], dbf[t + sy + 2 * vs],
                      dbf[t + sy + 3 * vs], dbf[t + sz + 4 * vs], dbf[t + sz + 5 * vs]};
    double Ex[4] = {E[t], E[t + sz], E[t + sy], E[t + sy + sz]};
    double Ey[4] = {E[t + vs], E[t + sz + vs], E[t + sx + vs], E[t + sx + sz + vs]};
    double Ez[4] = {E[t + 2 * vs], E[t + sy + 2 * vs], E[t + sx + 2 * vs], E[t + sx + sy + 2 * vs]};
    Half3D<double> h;
    half_step_3d(q, qPX, qMX, qPY, qMY, qPZ, qMZ, bfNb, d12, Ex, Ey, Ez, dtdx, dtdy, dtdz, p, h);
    // faces: [dir][side] -> slot 2*dir + (side>0): qp = slot even, qm = slot odd
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      store8(faces + (2 * d) * v8, p, t, face_state_3d(h, d, -1, p));
      store8(faces + (2 * d + 1) * v8, p, t, face_state_3d(h, d, +1, p));
    }
#pragma unroll
    for (int e = 0; e < 3; ++e)
#pragma unroll
      for (int c = 0; c < 4; ++c) store8(edges + (4 * e + c) * v8, p, t, edge_state_3d(h, e, c, p));
  }
}

// ---- 2D trace, range [gw-2, size-gw+1)^2, reference src/muscl/MHDRunFunctors2D.h:848-904 ---
__global__ void __launch_bounds__(kBlock) k_trace2d(const double *__restrict__ U,
                                                    const double *__restrict__ Q,
                                                    double *__restrict__ faces,
                                                    double *__restrict__ edges, DevParams p,
                                                    const double *ctrl) {
  if (!active(ctrl)) return;
  const double dt = ctrl[CTRL_DT];
  const double dtdx = dt / p.dx, dtdy = dt / p.dy;
  const long long total = (long long)p.isize * p.jsize;
  const long long vs = p.vstride, sy = p.isize;
  const long long v8 = 8 * vs;
  const int lo = p.gw - 2;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < total;
       t += (long long)gridDim.x * blockDim.x) {
    int i = (int)(t % p.isize), j = (int)(t / p.isize);
    if (!(i >= lo && i < p.isize - p.gw + 1 && j >= lo && j < p.jsize - p.gw + 1)) continue;
    St<double> qNb[3][3];
    double bf[4][4][3];
#pragma unroll
    for (int di = 0; di < 3; ++di)
#pragma unroll
      for (int dj = 0; dj < 3; ++dj) load8(Q, p, t + (di - 1) + (dj - 1) * sy, qNb[di][dj]);
#pragma unroll
    for (int di = 0; di < 4; ++di)
#pragma unroll
      for (int dj = 0; dj < 4; ++dj) {
        long long c = t + (di - 1) + (dj - 1) * sy;
        bf[di][dj][0] = U[c + IA * vs];
        bf[di][dj][1] = U[c + IB * vs];
        bf[di][dj][2] = U[c + IC * vs];
      }
    Trace2D<double> tr;
    trace_cell_2d(qNb, bf, dtdx, dtdy, p, tr);
#pragma unroll
    for (int d = 0; d < 2; ++d) {
      store8(faces + (2 * d) * v8, p, t, tr.qp[d]);
      store8(faces + (2 * d + 1) * v8, p, t, tr.qm[d]);
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) store8(edges + c * v8, p, t, tr.qe[c]);
  }
}

// ---- fluxes at x/y/z faces, reference src/muscl/MHDRunFunctors3D.h:1839-1902 (and 2D) -----
__global__ void __launch_bounds__(kBlock) k_fluxes(const double *__restrict__ faces,
                                                   double *__restrict__ F, DevParams p,
                                                   const double *ctrl) {
  if (!active(ctrl)) return;
  const long long total = (long long)p.isize * p.jsize * p.ksize;
  const long long vs = p.vstride, v8 = 8 * vs;
  const int g = p.gw;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < total;
       t += (long long)gridDim.x * blockDim.x) {
    int i, j, k;
    decompose(p, t, i, j, k);
    bool in = i >= g && i < p.isize - g + 1 && j >= g && j < p.jsize - g + 1;
    if (p.dim == 3) in = in && k >= g && k < p.ksize - g + 1;
    if (!in) continue;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      if (d >= p.dim) break;
      St<double> ql, qr, fl;
      const long long sd = d == 0 ? 1 : (d == 1 ? (long long)p.isize : p.plane);
      load8(faces + (2 * d + 1) * v8, p, t - sd, ql);  // qm of the left cell
      load8(faces + (2 * d) * v8, p, t, qr);               // qp of this cell
      if (d == 1) {
        double tmp = ql[IU]; ql[IU] = ql[IV]; ql[IV] = tmp;
        tmp = ql[IA]; ql[IA] = ql[IB]; ql[IB] = tmp;
        tmp = qr[IU]; qr[IU] = qr[IV]; qr[IV] = tmp;
        tmp = qr[IA]; qr[IA] = qr[IB]; qr[IB] = tmp;
      } else if (d == 2) {
        double tmp = ql[IU]; ql[IU] = ql[IW]; ql[IW] = tmp;
        tmp = ql[IA]; ql[IA] = ql[IC]; ql[IC] = tmp;
        tmp = qr[IU]; qr[IU] = qr[IW]; qr[IW] = tmp;
        tmp = qr[IA]; qr[IA] = qr[IC]; qr[IC] = tmp;
      }
      riemann_mhd(ql, qr, fl, p);
      store8(F + d * v8, p, t, fl);
    }
  }
}

// ---- EMFs at edges, reference src/muscl/MHDRunFunctors3D.h:2183-2230, 2D :654-680 ---------
__global__ void __launch_bounds__(kBlock) k_emf(const double *__restrict__ edges,
                                                double *__restrict__ Emf, DevParams p,
                                                const double *ctrl) {
  if (!active(ctrl)) return;
  const long long total = (long long)p.isize * p.jsize * p.ksize;
  const long long vs = p.vstride, v8 = 8 * vs;
  const long long sx = 1, sy = p.isize, sz = p.plane;
  const int g = p.gw;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < total;
       t += (long long)gridDim.x * blockDim.x) {
    int i, j, k;
    decompose(p, t, i, j, k);
    bool in = i >= g && i < p.isize - g + 1 && j >= g && j < p.jsize - g + 1;
    if (p.dim == 3) in = in && k >= g && k < p.ksize - g + 1;
    if (!in) continue;
    St<double> a, b, c, d;
    // edges layout: (4*e + corner) * v8, e = 0 X, 1 Y, 2 Z (2D uses e = 0 for its Z edges)
    const int ez = p.dim == 3 ? 2 : 0;
    load8(edges + (4 * ez + IRT) * v8, p, t - sx - sy, a);
    load8(edges + (4 * ez + IRB) * v8, p, t - sx, b);
    load8(edges + (4 * ez + ILT) * v8, p, t - sy, c);
    load8(edges + (4 * ez + ILB) * v8, p, t, d);
    Emf[t + I_EMFZ * vs] = compute_emf<EMFZ>(a, b, c, d, p);
    if (p.dim == 3) {
      // EMFY takes RB and LT swapped (reference :2212-2217)
      load8(edges + (4 * 1 + IRT) * v8, p, t - sx - sz, a);
      load8(edges + (4 * 1 + ILT) * v8, p, t - sz, b);
      load8(edges + (4 * 1 + IRB) * v8, p, t - sx, c);
      load8(edges + (4 * 1 + ILB) * v8, p, t, d);
      Emf[t + I_EMFY * vs] = compute_emf<EMFY>(a, b, c, d, p);
      load8(edges + (4 * 0 + IRT) * v8, p, t - sy - sz, a);
      load8(edges + (4 * 0 + IRB) * v8, p, t - sy, b);
      load8(edges + (4 * 0 + ILT) * v8, p, t - sz, c);
      load8(edges + (4 * 0 + ILB) * v8, p, t, d);
      Emf[t + I_EMFX * vs] = compute_emf<EMFX>(a, b, c, d, p);
    }
  }
}

// ---- flux update + CT (+ copy of ghost cells), references
// src/muscl/MHDRunFunctors{2D,3D}.h UpdateFunctor / UpdateEmfFunctor, applied in that order
// to data_out = copy of data_in (src/muscl/SolverMHDMuscl.cpp:385, 477) -----------------------
__global__ void __launch_bounds__(kBlock) k_update(const double *__restrict__ Uin,
                                                   double *__restrict__ Uout,
                                                   const double *__restrict__ F,
                                                   const double *__restrict__ Emf, DevParams p,
                                                   const double *ctrl) {
  if (!active(ctrl)) return;
  const double dt = ctrl[CTRL_DT];
  const double dtdx = dt / p.dx, dtdy = dt / p.dy, dtdz = p.dim == 3 ? dt / p.dz : 0.0;
  const long long total = (long long)p.isize * p.jsize * p.ksize;
  const long long vs = p.vstride, v8 = 8 * vs;
  const long long sx = 1, sy = p.isize, sz = p.plane;
  const int g = p.gw;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < total;
       t += (long long)gridDim.x * blockDim.x) {
    int i, j, k;
    decompose(p, t, i, j, k);
    bool in = i >= g && i < p.isize - g && j >= g && j < p.jsize - g;
    if (p.dim == 3) in = in && k >= g && k < p.ksize - g;
    St<double> u;
    load8(Uin, p, t, u);
    if (in) {
      St<double> f;
      const double *Fx = F, *Fy = F + v8, *Fz = F + 2 * v8;
      load8(Fx, p, t, f);
      u[ID] += f[ID] * dtdx; u[IP] += f[IP] * dtdx; u[IU] += f[IU] * dtdx;
      u[IV] += f[IV] * dtdx; u[IW] += f[IW] * dtdx;
      if (p.dim == 2) u[IC] += f[IC] * dtdx;
      load8(Fx, p, t + sx, f);
      u[ID] -= f[ID] * dtdx; u[IP] -= f[IP] * dtdx; u[IU] -= f[IU] * dtdx;
      u[IV] -= f[IV] * dtdx; u[IW] -= f[IW] * dtdx;
      if (p.dim == 2) u[IC] -= f[IC] * dtdx;
      load8(Fy, p, t, f);
      u[ID] += f[ID] * dtdy; u[IP] += f[IP] * dtdy; u[IU] += f[IV] * dtdy;
      u[IV] += f[IU] * dtdy; u[IW] += f[IW] * dtdy;
      if (p.dim == 2) u[IC] += f[IC] * dtdy;
      load8(Fy, p, t + sy, f);
      u[ID] -= f[ID] * dtdy; u[IP] -= f[IP] * dtdy; u[IU] -= f[IV] * dtdy;
      u[IV] -= f[IU] * dtdy; u[IW] -= f[IW] * dtdy;
      if (p.dim == 2) u[IC] -= f[IC] * dtdy;
      if (p.dim == 3) {
        load8(Fz, p, t, f);
        u[ID] += f[ID] * dtdz; u[IP] += f[IP] * dtdz; u[IU] += f[IW] * dtdz;
        u[IV] += f[IV] * dtdz; u[IW] += f[IU] * dtdz;
        load8(Fz, p, t + sz, f);
        u[ID] -= f[ID] * dtdz; u[IP] -= f[IP] * dtdz; u[IU] -= f[IW] * dtdz;
        u[IV] -= f[IV] * dtdz; u[IW] -= f[IU] * dtdz;
        // CT, reference src/muscl/MHDRunFunctors3D.h:2600-2620 (fixed order)
        auto em = [&](long long off, int c) { return Emf[t + off + c * vs]; };
        u[IA] += (em(sy, I_EMFZ) - em(0, I_EMFZ)) * dtdy;
        u[IB] -= (em(sx, I_EMFZ) - em(0, I_EMFZ)) * dtdx;
        u[IA] -= (em(sz, I_EMFY) - em(0, I_EMFY)) * dtdz;
        u[IB] += (em(sz, I_EMFX) - em(0, I_EMFX)) * dtdz;
        u[IC] += (em(sx, I_EMFY) - em(0, I_EMFY)) * dtdx;
        u[IC] -= (em(sy, I_EMFX) - em(0, I_EMFX)) * dtdy;
      } else {
        // CT, reference src/muscl/MHDRunFunctors2D.h:1690-1691
        u[IA] += (Emf[t + sy] - Emf[t]) * dtdy;
        u[IB] -= (Emf[t + sx] - Emf[t]) * dtdx;
      }
    }
    store8(Uout, p, t, u);
  }
}

// ---- time-loop control (single thread) ------------------------------------------------------
// before a step: stop when SolverBase::finished() holds (src/shared/SolverBase.cpp:195-201)
// if the step was asked to check
__global__ void k_begin(double *ctrl, int check_finished) {
  if (threadIdx.x != 0) return;
  if (check_finished) {
    double t = ctrl[CTRL_T], tEnd = ctrl[CTRL_TEND];
    bool done = t >= (tEnd - 1e-14) || ctrl[CTRL_ITER] >= ctrl[CTRL_NSTEPMAX];
    ctrl[CTRL_ACTIVE] = done ? 0.0 : 1.0;
  } else {
    ctrl[CTRL_ACTIVE] = 1.0;
  }
}
// dt from the CFL partial maxima or the fixed value, clamped to tEnd - t,
// src/muscl/SolverMHDMuscl.h:737 and src/shared/SolverBase.cpp:173-177.  One block.
constexpr int kSetDtBlock = 1024;
__global__ void __launch_bounds__(kSetDtBlock) k_set_dt(double *ctrl,
                                                        const double *__restrict__ parts,
                                                        int nparts, double fixed_dt, int mode,
                                                        double cfl) {
  __shared__ double red[kSetDtBlock / 64];
  if (ctrl[CTRL_ACTIVE] == 0.0) return;
  if (mode == SETDT_FIXED) {
    // a fixed dt is clamped to tEnd like the CFL one (SolverBase.cpp:173-177), so a run of
    // fixed steps stops exactly at tEnd
    if (threadIdx.x == 0) {
      double dt = fixed_dt;
      if (ctrl[CTRL_T] + dt > ctrl[CTRL_TEND]) dt = ctrl[CTRL_TEND] - ctrl[CTRL_T];
      ctrl[CTRL_DT] = dt;
    }
    return;
  }
  // eight independent loads in flight per thread: the partials can number ~10^5
  double m[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  int b = threadIdx.x;
  for (; b + 7 * kSetDtBlock < nparts; b += 8 * kSetDtBlock) {
#pragma unroll
    for (int u = 0; u < 8; ++u) m[u] = fmax(m[u], parts[b + u * kSetDtBlock]);
  }
  for (; b < nparts; b += kSetDtBlock) m[0] = fmax(m[0], parts[b]);
#pragma unroll
  for (int u = 1; u < 8; ++u) m[0] = fmax(m[0], m[u]);
  const double mx = block_max(m[0], red);
  if (threadIdx.x != 0) return;
  ctrl[CTRL_INVDT] = mx;
  if (mode == SETDT_CFL) {
    double dt = cfl / mx;
    if (ctrl[CTRL_T] + dt > ctrl[CTRL_TEND]) dt = ctrl[CTRL_TEND] - ctrl[CTRL_T];
    ctrl[CTRL_DT] = dt;
  }
}
// after a step: ++iteration; t += dt (src/shared/SolverBase.cpp:216-218)
__global__ void k_end(double *ctrl) {
  if (threadIdx.x != 0 || ctrl[CTRL_ACTIVE] == 0.0) return;
  ctrl[CTRL_ITER] += 1.0;
  ctrl[CTRL_T] += ctrl[CTRL_DT];
}

int grid_for(long long n) {
  long long g = (n + kBlock - 1) / kBlock;
  if (g > 65536) g = 65536;
  if (g < 1) g = 1;
  return (int)g;
}

}  // namespace

// -------------------------------------------------------------------------------------------
// launchers
// -------------------------------------------------------------------------------------------
hipError_t staged_boundaries(const DevParams &p, double *U, double *copy, const double *ctrl,
                             hipStream_t s, StageTimer *tm) {
  if (tm) tm->mark(ST_BOUNDARIES, true);
  hipLaunchKernelGGL(k_boundaries, dim3(grid_for(ghost_shell_cells(p))), dim3(kBlock), 0, s, U,
                     copy, p, ctrl);
  if (tm) tm->mark(ST_BOUNDARIES, false);
  return hipGetLastError();
}

hipError_t staged_prim_dt(const DevParams &p, const double *U, double *Q, double *parts,
                          const double *ctrl, int do_dt, hipStream_t s, StageTimer *tm) {
  static_assert(kBlock == kCflBlock, "CFL partials assume kCflBlock threads per block");
  PPK_TIMED(tm, ST_PRIM_DT,
            hipLaunchKernelGGL(k_prim_dt, dim3(cfl_blocks(p)), dim3(kBlock), 0, s, U, Q, p, parts,
                               ctrl, do_dt));
  return hipGetLastError();
}

hipError_t staged_begin(double *ctrl, int check_finished, hipStream_t s) {
  hipLaunchKernelGGL(k_begin, dim3(1), dim3(64), 0, s, ctrl, check_finished);
  return hipGetLastError();
}
hipError_t staged_set_dt(double *ctrl, const double *parts, int nparts, double fixed_dt, int mode,
                         double cfl, hipStream_t s) {
  hipLaunchKernelGGL(k_set_dt, dim3(1), dim3(kSetDtBlock), 0, s, ctrl, parts, nparts, fixed_dt,
                     mode, cfl);
  return hipGetLastError();
}
hipError_t staged_end(double *ctrl, hipStream_t s) {
  hipLaunchKernelGGL(k_end, dim3(1), dim3(64), 0, s, ctrl);
  return hipGetLastError();
}

hipError_t staged_godunov(const DevParams &p, const double *Uin, double *Uout, const double *Q,
                          StagedScratch &w, const double *ctrl, hipStream_t s, StageTimer *tm) {
  long long n = (long long)p.isize * p.jsize * p.ksize;
  dim3 g(grid_for(n)), b(kBlock);
  if (p.dim == 3) {
    PPK_TIMED(tm, ST_ELEC, hipLaunchKernelGGL(k_elec3d, g, b, 0, s, Uin, Q, w.elec, p, ctrl));
    PPK_TIMED(tm, ST_MAGSLOPES, hipLaunchKernelGGL(k_magslopes3d, g, b, 0, s, Uin, w.dbf, p, ctrl));
    PPK_TIMED(tm, ST_TRACE, hipLaunchKernelGGL(k_trace3d, g, b, 0, s, Uin, Q, w.elec, w.dbf, w.faces,
                                               w.edges, p, ctrl));
  } else {
    PPK_TIMED(tm, ST_TRACE,
              hipLaunchKernelGGL(k_trace2d, g, b, 0, s, Uin, Q, w.faces, w.edges, p, ctrl));
  }
  PPK_TIMED(tm, ST_FLUXES, hipLaunchKernelGGL(k_fluxes, g, b, 0, s, w.faces, w.flux, p, ctrl));
  PPK_TIMED(tm, ST_EMF, hipLaunchKernelGGL(k_emf, g, b, 0, s, w.edges, w.emf, p, ctrl));
  PPK_TIMED(tm, ST_UPDATE,
            hipLaunchKernelGGL(k_update, g, b, 0, s, Uin, Uout, w.flux, w.emf, p, ctrl));
  return hipGetLastError();
}

}  // namespace PPK_NS
}  // namespace ppk
