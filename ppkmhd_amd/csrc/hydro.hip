// Euler (hydro) MUSCL-Hancock step of the reference's Hydro_Muscl_2D / Hydro_Muscl_3D,
// implementation version 0 (src/muscl/SolverHydroMuscl.cpp:257-405):
//   ghost fill -> primitives -> fluxes of the left faces of [gw, size-gw] -> update.
// Two forms:
//  - staged: one kernel per reference stage, primitives and fluxes in HBM (the structural twin
//    of the reference, kept as the yardstick);
//  - fused: one marching kernel per step.  A wave owns 64 consecutive x columns (60 of them
//    output) and walks along y (2D) or z (3D, with 8 waves of a block stacked along y and
//    exchanging rows through LDS).  The primitives, slopes and Hancock sources of a cell are
//    computed once and kept in registers while the march passes it, each face is solved once,
//    the update of a cell is finished one row / plane later (its upper flux), and the CFL
//    maximum of the new state is folded into the same pass.  Per cell HBM traffic is one read
//    and one write of the state.
// Every face state, flux and update uses the reference's expressions in its order, so the
// strict build (-ffp-contract=off) gives the reference's double results bit for bit.
//
// This file is compiled twice: PPK_NS=fast (FMA contraction allowed) and PPK_NS=strict.
#include <hip/hip_runtime.h>

#include "hydro_kernels.hpp"

namespace ppk {
namespace PPK_NS {

namespace {

using hyd::HSt;
constexpr int kBlock = 256;

__device__ __forceinline__ bool active(const double *ctrl) { return ctrl[CTRL_ACTIVE] != 0.0; }

int grid_for(long long n) {
  long long g = (n + kBlock - 1) / kBlock;
  if (g > 65536) g = 65536;
  if (g < 1) g = 1;
  return (int)g;
}

template <int NV>
__device__ __forceinline__ void load(const double *__restrict__ a, long long vs, long long c,
                                     HSt<NV> &s) {
#pragma unroll
  for (int v = 0; v < NV; ++v) s[v] = a[c + v * vs];
}

template <int NV>
__device__ __forceinline__ void swap_normal(HSt<NV> &q, int dir) {
  if (dir == 0) return;
  const double t = q[IU];
  q[IU] = q[IU + dir];
  q[IU + dir] = t;
}

// Gravity terms take a mode GR: 0 none, 1 static gravity, -1 decided at run time by h.grav
// (the staged kernels); the fused kernels are compiled for GR = 0 and 1, so the step without
// gravity carries none of its code.
template <int GR>
__device__ __forceinline__ bool gravity_on(const HydroDev &h) {
  return GR == 1 || (GR < 0 && h.grav);
}

// the gravity field of padded-box cell c (zero without gravity, nothing loaded)
struct G3 {
  double x, y, z;
};
template <int GR, int NV>
__device__ __forceinline__ G3 cell_g(const HydroDev &h, long long c) {
  G3 g{0.0, 0.0, 0.0};
  if (gravity_on<GR>(h)) {
    g.x = h.gf[c];
    g.y = h.gf[c + h.vstride];
    if constexpr (NV == 5) g.z = h.gf[c + 2 * h.vstride];
  }
  return g;
}

// gravity predictor of a face state of a cell with field g, half a time step
// (HydroRunFunctors2D.h:650-660, 3D.h:485-497): v += (0.5 dt) g, before the normal is
// rotated into IU
template <int GR, int NV>
__device__ __forceinline__ void gravity_predict(HSt<NV> &q, const G3 &g, const HydroDev &h,
                                                double hdt) {
  if (!gravity_on<GR>(h)) return;
  q[IU] += hdt * g.x;
  q[IV] += hdt * g.y;
  if constexpr (NV == 5) q[IW] += hdt * g.z;
}

// GravitySourceTermFunctor2D/3D (HydroRunFunctors2D.h:1299-1331, 3D.h:1463-1500) on an updated
// cell o whose density before the step was rho_old: momentum += (0.5 dt) g (rho_old + rho_new),
// energy += the change of kinetic energy.  The 2D functor floors the new density at smallr,
// the 3D one does not.
template <int GR, int NV>
__device__ __forceinline__ void gravity_source(HSt<NV> &o, double rho_old, const G3 &g,
                                               const HydroDev &h, double hdt) {
  if (!gravity_on<GR>(h)) return;
  const double rho_new = NV == 4 ? fmax(h.P.smallr, o[ID]) : o[ID];
  const double ir = recip(rho_new);
  double mu = o[IU], mv = o[IV], mw = NV == 5 ? o[IW] : 0.0;
  const double ek_old = NV == 5 ? divr(0.5 * (mu * mu + mv * mv + mw * mw), rho_new, ir)
                                : divr(0.5 * (mu * mu + mv * mv), rho_new, ir);
  mu += hdt * g.x * (rho_old + rho_new);
  mv += hdt * g.y * (rho_old + rho_new);
  if constexpr (NV == 5) mw += hdt * g.z * (rho_old + rho_new);
  o[IU] = mu;
  o[IV] = mv;
  if constexpr (NV == 5) o[IW] = mw;
  const double ek_new = NV == 5 ? divr(0.5 * (mu * mu + mv * mv + mw * mw), rho_new, ir)
                                : divr(0.5 * (mu * mu + mv * mv), rho_new, ir);
  o[IP] += (ek_new - ek_old);
}

// CFL term of one cell (ComputeDtFunctor2D/3D, HydroRunFunctors2D.h:46-76, 3D.h:46-81); with
// gravity ComputeDtGravityFunctor2D/3D (HydroRunFunctors2D.h:120-170, 3D.h:130-185): the
// summed speeds u scaled by k / (sqrt(1 + 2k) - 1), k = |g|_1 cfl dx / u^2 floored at 1e-4,
// which makes g dt^2 / (2 dx) + u dt / dx = cfl, over the smallest cell size
template <int GR, int NV>
__device__ __forceinline__ double inv_dt_prim(const HSt<NV> &q, double c, const G3 &g,
                                              const HydroDev &h) {
  if (gravity_on<GR>(h)) {
    double vel = 0.0;
    vel += c + fabs(q[IU]);
    vel += c + fabs(q[IV]);
    if constexpr (NV == 5) vel += c + fabs(q[IW]);
    double k = NV == 5 ? fabs(g.x) + fabs(g.y) + fabs(g.z) : fabs(g.x) + fabs(g.y);
    const double v2 = vel * vel;
    k = k * divr(h.cfl * h.dxmin, v2, recip(v2));
    k = fmax(k, 1e-4);
    const double den = msqrt(1.0 + 2.0 * k) - 1.0;
    vel = vel * divr(k, den, recip(den));
    return divr(vel, h.dxmin, h.idxmin);
  }
  const double vx = c + fabs(q[IU]);
  const double vy = c + fabs(q[IV]);
  if constexpr (NV == 5) {
    const double vz = c + fabs(q[IW]);
    return divr(vx, h.dx, h.idx) + divr(vy, h.dy, h.idy) + divr(vz, h.dz, h.idz);
  } else {
    return divr(vx, h.dx, h.idx) + divr(vy, h.dy, h.idy);
  }
}
template <int GR, int NV>
__device__ __forceinline__ double inv_dt_cell(const HSt<NV> &u, const G3 &g, const HydroDev &h) {
  HSt<NV> q;
  double c;
  hyd::prims(u, h.P, c, q);
  return inv_dt_prim<GR>(q, c, g, h);
}

// ---- ghost fill, one direction, both faces (MakeBoundariesFunctor2D/3D,
// src/shared/BoundariesFunctors.h:18-531): Dirichlet reflects and flips the normal velocity,
// Neumann copies the first interior cell, periodic wraps
template <int NV>
__global__ void __launch_bounds__(kBlock) k_hbnd(double *__restrict__ U, double *__restrict__ C,
                                                 HydroDev h, int dir, const double *ctrl) {
  if (!active(ctrl)) return;
  const int gw = h.gw;
  const int n = dir == 0 ? h.nx : (dir == 1 ? h.ny : h.nz);
  const int s1 = dir == 0 ? h.jsize : h.isize;
  const int s2 = dir == 2 ? h.jsize : h.ksize;
  const long long total = 2LL * gw * s1 * s2;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < total;
       t += (long long)gridDim.x * blockDim.x) {
    const int g = (int)(t % gw);
    long long r = t / gw;
    const int t1 = (int)(r % s1);
    r /= s1;
    const int t2 = (int)(r % s2);
    const int upper = (int)(r / s2);
    const int bc = h.bc[2 * dir + upper];
    const int c = upper ? g + n + gw : g;
    int c0;
    if (bc == PPK_BC_DIRICHLET)
      c0 = upper ? 2 * n + 2 * gw - 1 - c : 2 * gw - 1 - c;
    else if (bc == PPK_BC_NEUMANN)
      c0 = upper ? n + gw - 1 : gw;
    else
      c0 = upper ? c - n : n + c;
    int i, j, k, i0, j0, k0;
    if (dir == 0) { i = c; j = t1; k = t2; i0 = c0; j0 = j; k0 = k; }
    else if (dir == 1) { i = t1; j = c; k = t2; i0 = i; j0 = c0; k0 = k; }
    else { i = t1; j = t2; k = c; i0 = i; j0 = j; k0 = c0; }
    const long long dst = ((long long)k * h.jsize + j) * h.isize + i;
    const long long src = ((long long)k0 * h.jsize + j0) * h.isize + i0;
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      const double sign = (bc == PPK_BC_DIRICHLET && v == IU + dir) ? -1.0 : 1.0;
      const double x = U[src + v * h.vstride] * sign;
      U[dst + v * h.vstride] = x;
      if (C) C[dst + v * h.vstride] = x;
    }
  }
}

// ---- primitives of every cell (ConvertToPrimitivesFunctor, SolverHydroMuscl.h) and the CFL
// maximum over the interior, one partial per block
template <int NV>
__global__ void __launch_bounds__(kBlock) k_hprim(const double *__restrict__ U,
                                                  double *__restrict__ Q, HydroDev h,
                                                  double *__restrict__ parts, const double *ctrl,
                                                  int do_dt) {
  if (!active(ctrl)) return;
  __shared__ double red[kBlock / 64];
  const long long total = (long long)h.isize * h.jsize * h.ksize;
  double best = 0.0;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < total;
       t += (long long)gridDim.x * blockDim.x) {
    HSt<NV> u, q;
    load(U, h.vstride, t, u);
    double c;
    hyd::prims(u, h.P, c, q);
    if (Q) {
#pragma unroll
      for (int v = 0; v < NV; ++v) Q[t + v * h.vstride] = q[v];
    }
    if (do_dt) {
      const int i = (int)(t % h.isize);
      const long long r = t / h.isize;
      const int j = (int)(r % h.jsize);
      const int k = (int)(r / h.jsize);
      const int g = h.gw;
      const bool inner = i >= g && i < h.isize - g && j >= g && j < h.jsize - g &&
                         (NV == 4 || (k >= g && k < h.ksize - g));
      if (inner) best = fmax(best, inv_dt_prim<-1>(q, c, cell_g<-1, NV>(h, t), h));
    }
  }
  if (!do_dt) return;
  best = block_max(best, red);
  if (threadIdx.x == 0) parts[blockIdx.x] = best;
}

// ---- staged fluxes (ComputeAndStoreFluxesFunctor2D/3D, HydroRunFunctors2D.h:490-750,
// HydroRunFunctors3D.h:300-720): for every cell of [gw, size-gw] the flux through its left
// face in each direction, slopes of the cell and of its left neighbour recomputed from Q,
// the normal velocity rotated into IU for the solver and back, times dt/dx
template <int DIM, int NV>
__device__ __forceinline__ void cell_slopes(const double *__restrict__ Q, const HydroDev &h,
                                            long long c, HSt<NV> &q, HSt<NV> &dX, HSt<NV> &dY,
                                            HSt<NV> &dZ) {
  HSt<NV> qpx, qmx, qpy, qmy, qpz, qmz;
  load(Q, h.vstride, c, q);
  load(Q, h.vstride, c + 1, qpx);
  load(Q, h.vstride, c - 1, qmx);
  load(Q, h.vstride, c + h.isize, qpy);
  load(Q, h.vstride, c - h.isize, qmy);
  if constexpr (DIM == 3) {
    load(Q, h.vstride, c + h.plane, qpz);
    load(Q, h.vstride, c - h.plane, qmz);
  } else {
    qpz = q;
    qmz = q;
  }
  hyd::slopes(q, qpx, qmx, qpy, qmy, qpz, qmz, h.P, dX, dY, dZ);
}

template <int DIM, int NV, int RS>
__global__ void __launch_bounds__(kBlock) k_hflux(const double *__restrict__ Q,
                                                  double *__restrict__ F, HydroDev h,
                                                  const double *ctrl) {
  if (!active(ctrl)) return;
  const double dt = ctrl[CTRL_DT];
  const double dtd[3] = {dt / h.dx, dt / h.dy, DIM == 3 ? dt / h.dz : 0.0};
  const double hdt = 0.5 * dt;
  const int nsx = h.nx + 1, nsy = h.ny + 1, nsz = DIM == 3 ? h.nz + 1 : 1;
  const long long total = (long long)nsx * nsy * nsz;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < total;
       t += (long long)gridDim.x * blockDim.x) {
    const int i = h.gw + (int)(t % nsx);
    const long long r = t / nsx;
    const int j = h.gw + (int)(r % nsy);
    const int k = DIM == 3 ? h.gw + (int)(r / nsy) : 0;
    const long long c = ((long long)k * h.jsize + j) * h.isize + i;
    HSt<NV> q, dX, dY, dZ, s0;
    cell_slopes<DIM>(Q, h, c, q, dX, dY, dZ);
    hyd::sources(q, dX, dY, dZ, h.P.gamma0, s0);
#pragma unroll 1
    for (int d = 0; d < DIM; ++d) {
      const long long st = d == 0 ? 1 : (d == 1 ? (long long)h.isize : h.plane);
      HSt<NV> qn, nX, nY, nZ, ns0;
      cell_slopes<DIM>(Q, h, c - st, qn, nX, nY, nZ);
      hyd::sources(qn, nX, nY, nZ, h.P.gamma0, ns0);
      HSt<NV> qr, ql, f;
      hyd::face_state(q, d == 0 ? dX : (d == 1 ? dY : dZ), s0, dtd[d], false, h.P.smallr, qr);
      hyd::face_state(qn, d == 0 ? nX : (d == 1 ? nY : nZ), ns0, dtd[d], true, h.P.smallr, ql);
      gravity_predict<-1>(ql, cell_g<-1, NV>(h, c - st), h, hdt);
      gravity_predict<-1>(qr, cell_g<-1, NV>(h, c), h, hdt);
      swap_normal(ql, d);
      swap_normal(qr, d);
      hyd::riemann_rs<RS>(ql, qr, f, h.P);
      swap_normal(f, d);
#pragma unroll
      for (int v = 0; v < NV; ++v) F[(long long)(d * NV + v) * h.vstride + c] = f[v] * dtd[d];
    }
  }
}

// ---- staged update (UpdateFunctor2D/3D, HydroRunFunctors2D.h:776-840, 3D.h:740-815)
template <int DIM, int NV>
__global__ void __launch_bounds__(kBlock) k_hupdate(double *__restrict__ U,
                                                    const double *__restrict__ F, HydroDev h,
                                                    const double *ctrl) {
  if (!active(ctrl)) return;
  const int nz = DIM == 3 ? h.nz : 1;
  const long long total = (long long)h.nx * h.ny * nz;
  const long long vs = h.vstride;
  const double hdt = 0.5 * ctrl[CTRL_DT];
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < total;
       t += (long long)gridDim.x * blockDim.x) {
    const int i = h.gw + (int)(t % h.nx);
    const long long r = t / h.nx;
    const int j = h.gw + (int)(r % h.ny);
    const int k = DIM == 3 ? h.gw + (int)(r / h.ny) : 0;
    const long long c = ((long long)k * h.jsize + j) * h.isize + i;
    const double rho_old = U[c];  // U holds a copy of the step's input state
    HSt<NV> on;
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      const double *fx = F + (long long)v * vs;
      const double *fy = F + (long long)(NV + v) * vs;
      double o = U[c + v * vs];
      o += fx[c];
      o -= fx[c + 1];
      o += fy[c];
      o -= fy[c + h.isize];
      if constexpr (DIM == 3) {
        const double *fz = F + (long long)(2 * NV + v) * vs;
        o += fz[c];
        o -= fz[c + h.plane];
      }
      on[v] = o;
    }
    gravity_source<-1>(on, rho_old, cell_g<-1, NV>(h, c), h, hdt);
#pragma unroll
    for (int v = 0; v < NV; ++v) U[c + v * vs] = on[v];
  }
}

// ---- fused 2D: one wave per strip of 60 output columns x strip_rows output rows ----------
// Lane l holds column i = gw + 60 * sx + l - 2; lanes 2..61 are the strip's output.  The wave
// walks the slope rows s = r0-1 .. r1: at row s it has Q(s-1), Q(s), Q(s+1) in registers, the
// x neighbours' Q from lanes l -/+ 1, computes slopes and sources of (i, s), solves the y face
// below row s (upper state of row s-1 kept from the previous row), the x face left of (i, s)
// (left state from lane l-1), and finishes the update of (i, s-1) with the flux of that y face.
constexpr int kWaves2 = 4;

template <int NV, int RS, int MINW, int GR>
__global__ void __launch_bounds__(kWaves2 * 64, MINW) k_hfused2d(const double *__restrict__ Uin,
                                                           double *__restrict__ Uout, HydroDev h,
                                                           double *__restrict__ parts,
                                                           const double *ctrl, int fold_cfl) {
  if (!active(ctrl)) return;
  const int lane = threadIdx.x & 63;
  const int wave = blockIdx.x * kWaves2 + (threadIdx.x >> 6);
  const int nsx = (h.nx + kHydroCols - 1) / kHydroCols;
  const int H = h.strip_rows;
  const int nsy = (h.ny + H - 1) / H;
  if (wave >= nsx * nsy) return;  // whole waves only
  const int sx = wave % nsx, sy = wave / nsx;
  const int i = h.gw + sx * kHydroCols + lane - 2;
  const int ic = i < h.isize ? i : h.isize - 1;  // lanes past the box: clamped, never output
  const int r0 = h.gw + sy * H;
  const int r1 = min(r0 + H, h.gw + h.ny);
  const bool out_lane = lane >= 2 && lane < 2 + kHydroCols && i < h.gw + h.nx;
  const long long vs = h.vstride;
  const double dt = ctrl[CTRL_DT];
  const double dtdx = dt / h.dx, dtdy = dt / h.dy, hdt = 0.5 * dt;
  const hyd::HydroPar &P = h.P;

  HSt<NV> qm, q0, u0, t;
  double c, rho_prev = 0.0;  // density of (i, s-1) before the step, for the gravity source
  load(Uin, vs, (long long)(r0 - 2) * h.isize + ic, t);
  hyd::prims(t, P, c, qm);
  load(Uin, vs, (long long)(r0 - 1) * h.isize + ic, u0);
  hyd::prims(u0, P, c, q0);
  // gravity field of the row the face states are built for, and of the row being finished
  G3 g0 = cell_g<GR, NV>(h, (long long)(r0 - 1) * h.isize + ic), g_prev = g0;
  HSt<NV> ymax_prev, o_prev;
  double best = 0.0;
#pragma unroll 1
  for (int s = r0 - 1; s <= r1; ++s) {
    HSt<NV> un, qp;
    load(Uin, vs, (long long)(s + 1) * h.isize + ic, un);
    hyd::prims(un, P, c, qp);
    const G3 gn = cell_g<GR, NV>(h, (long long)(s + 1) * h.isize + ic);
    HSt<NV> qxm, qxp, dX, dY, dZ, s0;
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      qxm[v] = __shfl_up(q0[v], 1, 64);
      qxp[v] = __shfl_down(q0[v], 1, 64);
    }
    hyd::slopes(q0, qxp, qxm, qp, qm, q0, q0, P, dX, dY, dZ);
    hyd::sources(q0, dX, dY, dZ, P.gamma0, s0);
    if (s >= r0) {
      HSt<NV> ql = ymax_prev, qr, fy;
      hyd::face_state(q0, dY, s0, dtdy, false, P.smallr, qr);
      gravity_predict<GR>(qr, g0, h, hdt);
      swap_normal(ql, 1);
      swap_normal(qr, 1);
      hyd::riemann_rs<RS>(ql, qr, fy, P);
      swap_normal(fy, 1);
#pragma unroll
      for (int v = 0; v < NV; ++v) fy[v] = fy[v] * dtdy;
      if (s > r0 && out_lane) {  // finish (i, s-1): its upper y face is this one
        const long long cp = (long long)(s - 1) * h.isize + i;
#pragma unroll
        for (int v = 0; v < NV; ++v) o_prev[v] -= fy[v];
        gravity_source<GR>(o_prev, rho_prev, g_prev, h, hdt);
#pragma unroll
        for (int v = 0; v < NV; ++v) Uout[cp + v * vs] = o_prev[v];
        if (fold_cfl) best = fmax(best, inv_dt_cell<GR>(o_prev, g_prev, h));
      }
      if (s < r1) {
        HSt<NV> xmin, xmax, xl, fx;
        hyd::face_state(q0, dX, s0, dtdx, false, P.smallr, xmin);
        hyd::face_state(q0, dX, s0, dtdx, true, P.smallr, xmax);
        gravity_predict<GR>(xmin, g0, h, hdt);
        gravity_predict<GR>(xmax, g0, h, hdt);
#pragma unroll
        for (int v = 0; v < NV; ++v) xl[v] = __shfl_up(xmax[v], 1, 64);
        hyd::riemann_rs<RS>(xl, xmin, fx, P);
#pragma unroll
        for (int v = 0; v < NV; ++v) {
          fx[v] = fx[v] * dtdx;
          const double fxr = __shfl_down(fx[v], 1, 64);
          double o = u0[v];
          o += fx[v];
          o -= fxr;
          o += fy[v];
          o_prev[v] = o;
        }
        rho_prev = u0[ID];
        g_prev = g0;
      }
    }
    hyd::face_state(q0, dY, s0, dtdy, true, P.smallr, ymax_prev);
    gravity_predict<GR>(ymax_prev, g0, h, hdt);
    qm = q0;
    q0 = qp;
    u0 = un;
    g0 = gn;
  }
  if (fold_cfl) {
    for (int off = 32; off > 0; off >>= 1) best = fmax(best, __shfl_xor(best, off, 64));
    if (lane == 0) parts[wave] = best;
  }
}

// ---- fused 3D: a block of 8 waves, 16 rows of 32 columns stacked along y, marches along z --
// Thread t holds row r = t / 32 (two rows per wave), j = gw + 12 * by + r - 2, and column
// c = t % 32, i = gw + 28 * bx + c - 2; rows 2..13 and columns 2..29 are the block's output
// (32-wide rows: 336 output cells per block and plane, against 240 with one 64-wide row per
// wave).  x neighbours come from LDS and shuffles within the 32 lanes of a row.  At slope
// plane s each thread keeps Q(s-1), Q(s),
// Q(s+1) of its column in registers; plane s of Q goes through LDS for the x and y
// neighbours, the upper y-face states of each row through LDS to the row above, and the y
// fluxes back through LDS to the row below.  The z face below plane s uses the upper z state
// of plane s-1 kept in registers, so the update of (i, j, s-1) completes at plane s.
constexpr int kW3 = kHydroWaves3;

template <int NV, int RS, int MINW, int GR>
__global__ void __launch_bounds__(kW3 * 64, MINW) k_hfused3d(const double *__restrict__ Uin,
                                                       double *__restrict__ Uout, HydroDev h,
                                                       double *__restrict__ parts,
                                                       const double *ctrl, int fold_cfl) {
  constexpr int RW = kHydroRowW3, NR = kHydroBlockRows3;
  __shared__ double sQ[2][NV][NR][RW];  // Q of plane s (parity s & 1), later the y fluxes
  __shared__ double sT[NV][NR][RW];     // upper y-face states of plane s
  __shared__ double red[kW3];
  if (!active(ctrl)) return;
  const int lane = threadIdx.x % RW;  // column within the row
  const int ly = threadIdx.x / RW;    // row of the block
  const int nsx = (h.nx + kHydroCols3 - 1) / kHydroCols3;
  const int nsy = (h.ny + kHydroRows3 - 1) / kHydroRows3;
  const int KH = h.strip_rows;
  const int bx = blockIdx.x % nsx;
  const int by = (blockIdx.x / nsx) % nsy;
  const int bz = blockIdx.x / (nsx * nsy);
  const int i = h.gw + bx * kHydroCols3 + lane - 2;
  const int j = h.gw + by * kHydroRows3 + ly - 2;
  const int ic = i < h.isize ? i : h.isize - 1;  // clamped halo, never output
  const int jc = j < h.jsize ? j : h.jsize - 1;
  const int k0 = h.gw + bz * KH;
  const int k1 = min(k0 + KH, h.gw + h.nz);
  const bool slope_row = ly >= 1 && ly <= NR - 2;
  const bool yface_row = ly >= 2 && ly <= NR - 2;
  const bool out_row = ly >= 2 && ly < 2 + kHydroRows3;
  const bool out = out_row && lane >= 2 && lane < 2 + kHydroCols3 && i < h.gw + h.nx &&
                   j < h.gw + h.ny;
  const long long vs = h.vstride;
  const long long col = (long long)jc * h.isize + ic;
  const double dt = ctrl[CTRL_DT];
  const double dtdx = dt / h.dx, dtdy = dt / h.dy, dtdz = dt / h.dz, hdt = 0.5 * dt;
  const hyd::HydroPar &P = h.P;
  const int lm = lane > 0 ? lane - 1 : 0, lp = lane < RW - 1 ? lane + 1 : RW - 1;

  HSt<NV> qm, q0, u0, t;
  double c, rho_prev = 0.0;  // density of (i, j, s-1) before the step, for the gravity source
  load(Uin, vs, (long long)(k0 - 2) * h.plane + col, t);
  hyd::prims(t, P, c, qm);
  load(Uin, vs, (long long)(k0 - 1) * h.plane + col, u0);
  hyd::prims(u0, P, c, q0);
  // gravity field of the plane the face states are built for, and of the plane being finished
  G3 g0 = cell_g<GR, NV>(h, (long long)(k0 - 1) * h.plane + col), g_prev = g0;
  HSt<NV> zmax_prev, o_prev;
  double best = 0.0;
#pragma unroll 1
  for (int s = k0 - 1; s <= k1; ++s) {
    const int b = s & 1;
    HSt<NV> un, qp;
    load(Uin, vs, (long long)(s + 1) * h.plane + col, un);
    hyd::prims(un, P, c, qp);
    const G3 gn = cell_g<GR, NV>(h, (long long)(s + 1) * h.plane + col);
#pragma unroll
    for (int v = 0; v < NV; ++v) sQ[b][v][ly][lane] = q0[v];
    __syncthreads();
    HSt<NV> dX, dY, dZ, s0;
    if (slope_row) {
      HSt<NV> qxm, qxp, qym, qyp;
#pragma unroll
      for (int v = 0; v < NV; ++v) {
        qxm[v] = sQ[b][v][ly][lm];
        qxp[v] = sQ[b][v][ly][lp];
        qym[v] = sQ[b][v][ly - 1][lane];
        qyp[v] = sQ[b][v][ly + 1][lane];
      }
      hyd::slopes(q0, qxp, qxm, qyp, qym, qp, qm, P, dX, dY, dZ);
      hyd::sources(q0, dX, dY, dZ, P.gamma0, s0);
      HSt<NV> ymax;
      hyd::face_state(q0, dY, s0, dtdy, true, P.smallr, ymax);
      gravity_predict<GR>(ymax, g0, h, hdt);
#pragma unroll
      for (int v = 0; v < NV; ++v) sT[v][ly][lane] = ymax[v];
    }
    __syncthreads();
    const bool plane_out = s >= k0 && s < k1;
    HSt<NV> fx, fz;
    if (s >= k0 && out_row) {  // z face below plane s (also the top face of plane k1 - 1)
      HSt<NV> ql = zmax_prev, qr;
      hyd::face_state(q0, dZ, s0, dtdz, false, P.smallr, qr);
      gravity_predict<GR>(qr, g0, h, hdt);
      swap_normal(ql, 2);
      swap_normal(qr, 2);
      hyd::riemann_rs<RS>(ql, qr, fz, P);
      swap_normal(fz, 2);
#pragma unroll
      for (int v = 0; v < NV; ++v) fz[v] = fz[v] * dtdz;
    }
    if (plane_out && yface_row) {  // y face below row j, flux into the Q slot of plane s
      HSt<NV> ql, qr, fy;
#pragma unroll
      for (int v = 0; v < NV; ++v) ql[v] = sT[v][ly - 1][lane];
      hyd::face_state(q0, dY, s0, dtdy, false, P.smallr, qr);
      gravity_predict<GR>(qr, g0, h, hdt);
      swap_normal(ql, 1);
      swap_normal(qr, 1);
      hyd::riemann_rs<RS>(ql, qr, fy, P);
      swap_normal(fy, 1);
#pragma unroll
      for (int v = 0; v < NV; ++v) sQ[b][v][ly][lane] = fy[v] * dtdy;
    }
    if (plane_out && out_row) {  // x face left of column i
      HSt<NV> xmin, xmax, xl;
      hyd::face_state(q0, dX, s0, dtdx, false, P.smallr, xmin);
      hyd::face_state(q0, dX, s0, dtdx, true, P.smallr, xmax);
      gravity_predict<GR>(xmin, g0, h, hdt);
      gravity_predict<GR>(xmax, g0, h, hdt);
#pragma unroll
      for (int v = 0; v < NV; ++v) xl[v] = __shfl_up(xmax[v], 1, RW);
      hyd::riemann_rs<RS>(xl, xmin, fx, P);
#pragma unroll
      for (int v = 0; v < NV; ++v) fx[v] = fx[v] * dtdx;
    }
    __syncthreads();
    if (out_row && s >= k0) {
      if (s > k0 && out) {  // finish (i, j, s-1): its upper z face is this one
        const long long cp = (long long)(s - 1) * h.plane + (long long)j * h.isize + i;
#pragma unroll
        for (int v = 0; v < NV; ++v) o_prev[v] -= fz[v];
        gravity_source<GR>(o_prev, rho_prev, g_prev, h, hdt);
#pragma unroll
        for (int v = 0; v < NV; ++v) Uout[cp + v * vs] = o_prev[v];
        if (fold_cfl) best = fmax(best, inv_dt_cell<GR>(o_prev, g_prev, h));
      }
      if (plane_out) {
#pragma unroll
        for (int v = 0; v < NV; ++v) {
          const double fxr = __shfl_down(fx[v], 1, RW);
          double o = u0[v];
          o += fx[v];
          o -= fxr;
          o += sQ[b][v][ly][lane];
          o -= sQ[b][v][ly + 1][lane];
          o += fz[v];
          o_prev[v] = o;
        }
        rho_prev = u0[ID];
        g_prev = g0;
      }
    }
    if (slope_row) {
      hyd::face_state(q0, dZ, s0, dtdz, true, P.smallr, zmax_prev);
      gravity_predict<GR>(zmax_prev, g0, h, hdt);
    }
    qm = q0;
    q0 = qp;
    u0 = un;
    g0 = gn;
  }
  if (fold_cfl) {
    best = block_max(best, red);
    if (threadIdx.x == 0) parts[blockIdx.x] = best;
  }
}

int fused2d_waves(const HydroDev &h) {
  const int nsx = (h.nx + kHydroCols - 1) / kHydroCols;
  const int nsy = (h.ny + h.strip_rows - 1) / h.strip_rows;
  return nsx * nsy;
}
int fused3d_blocks(const HydroDev &h) {
  const int nsx = (h.nx + kHydroCols3 - 1) / kHydroCols3;
  const int nsy = (h.ny + kHydroRows3 - 1) / kHydroRows3;
  const int nsz = (h.nz + h.strip_rows - 1) / h.strip_rows;
  return nsx * nsy * nsz;
}

}  // namespace

// -------------------------------------------------------------------------------------------
// launchers
// -------------------------------------------------------------------------------------------
hipError_t hydro_boundaries(const HydroDev &h, double *U, double *copy, const double *ctrl,
                            hipStream_t s) {
  for (int dir = 0; dir < h.dim; ++dir) {
    const int s1 = dir == 0 ? h.jsize : h.isize;
    const int s2 = dir == 2 ? h.jsize : h.ksize;
    const long long n = 2LL * h.gw * s1 * s2;
    if (h.dim == 3)
      hipLaunchKernelGGL(k_hbnd<5>, dim3(grid_for(n)), dim3(kBlock), 0, s, U, copy, h, dir, ctrl);
    else
      hipLaunchKernelGGL(k_hbnd<4>, dim3(grid_for(n)), dim3(kBlock), 0, s, U, copy, h, dir, ctrl);
  }
  return hipGetLastError();
}

int hydro_prim_parts(const HydroDev &h) {
  return grid_for((long long)h.isize * h.jsize * h.ksize) < 4096
             ? grid_for((long long)h.isize * h.jsize * h.ksize)
             : 4096;
}

hipError_t hydro_prim_dt(const HydroDev &h, const double *U, double *Q, double *parts,
                         const double *ctrl, int do_dt, hipStream_t s) {
  const dim3 g(hydro_prim_parts(h)), b(kBlock);
  if (h.dim == 3)
    hipLaunchKernelGGL(k_hprim<5>, g, b, 0, s, U, Q, h, parts, ctrl, do_dt);
  else
    hipLaunchKernelGGL(k_hprim<4>, g, b, 0, s, U, Q, h, parts, ctrl, do_dt);
  return hipGetLastError();
}

template <int RS>
void launch_staged_fluxes(const HydroDev &h, const double *Q, double *F, const double *ctrl,
                          hipStream_t s) {
  const dim3 b(kBlock);
  if (h.dim == 3) {
    const long long nf = (long long)(h.nx + 1) * (h.ny + 1) * (h.nz + 1);
    hipLaunchKernelGGL((k_hflux<3, 5, RS>), dim3(grid_for(nf)), b, 0, s, Q, F, h, ctrl);
  } else {
    const long long nf = (long long)(h.nx + 1) * (h.ny + 1);
    hipLaunchKernelGGL((k_hflux<2, 4, RS>), dim3(grid_for(nf)), b, 0, s, Q, F, h, ctrl);
  }
}

hipError_t hydro_staged_fluxes(const HydroDev &h, const double *Q, double *F,
                               const double *ctrl, hipStream_t s) {
  switch (h.P.riemann) {  // the solver as a compile-time constant of the kernel
    case PPK_RIEMANN_APPROX: launch_staged_fluxes<PPK_RIEMANN_APPROX>(h, Q, F, ctrl, s); break;
    case PPK_RIEMANN_HLL: launch_staged_fluxes<PPK_RIEMANN_HLL>(h, Q, F, ctrl, s); break;
    case PPK_RIEMANN_HLLC: launch_staged_fluxes<PPK_RIEMANN_HLLC>(h, Q, F, ctrl, s); break;
    default: launch_staged_fluxes<PPK_RIEMANN_LLF>(h, Q, F, ctrl, s); break;
  }
  return hipGetLastError();
}

hipError_t hydro_staged_update(const HydroDev &h, const double *F, double *Uout,
                               const double *ctrl, hipStream_t s) {
  const dim3 b(kBlock);
  const long long nu = (long long)h.nx * h.ny * (h.dim == 3 ? h.nz : 1);
  if (h.dim == 3)
    hipLaunchKernelGGL((k_hupdate<3, 5>), dim3(grid_for(nu)), b, 0, s, Uout, F, h, ctrl);
  else
    hipLaunchKernelGGL((k_hupdate<2, 4>), dim3(grid_for(nu)), b, 0, s, Uout, F, h, ctrl);
  return hipGetLastError();
}

int hydro_fused_parts(const HydroDev &h) {
  return h.dim == 3 ? fused3d_blocks(h) : fused2d_waves(h);
}

template <int RS, int GR>
void launch_fused_gr(const HydroDev &h, const double *Uin, double *Uout, double *parts,
                     const double *ctrl, int fold_cfl, hipStream_t s) {
  if (h.dim == 3) {
    if constexpr (RS != PPK_RIEMANN_HLL) {  // no 3D HLL (refused at create)
      const dim3 g(fused3d_blocks(h)), b(kW3 * 64);
      hipLaunchKernelGGL((k_hfused3d<5, RS, 1, GR>), g, b, 0, s, Uin, Uout, h, parts, ctrl,
                         fold_cfl);
    }
  } else {
    const int waves = fused2d_waves(h);
    const dim3 g((waves + kWaves2 - 1) / kWaves2), b(kWaves2 * 64);
    hipLaunchKernelGGL((k_hfused2d<4, RS, 1, GR>), g, b, 0, s, Uin, Uout, h, parts, ctrl,
                       fold_cfl);
  }
}

// the fused kernels with or without the gravity terms
template <int RS>
void launch_fused(const HydroDev &h, const double *Uin, double *Uout, double *parts,
                  const double *ctrl, int fold_cfl, hipStream_t s) {
  if (h.grav)
    launch_fused_gr<RS, 1>(h, Uin, Uout, parts, ctrl, fold_cfl, s);
  else
    launch_fused_gr<RS, 0>(h, Uin, Uout, parts, ctrl, fold_cfl, s);
}

hipError_t hydro_fused(const HydroDev &h, const double *Uin, double *Uout, double *parts,
                       const double *ctrl, int fold_cfl, hipStream_t s) {
  switch (h.P.riemann) {  // the solver as a compile-time constant of the kernel
    case PPK_RIEMANN_APPROX:
      launch_fused<PPK_RIEMANN_APPROX>(h, Uin, Uout, parts, ctrl, fold_cfl, s);
      break;
    case PPK_RIEMANN_HLL: launch_fused<PPK_RIEMANN_HLL>(h, Uin, Uout, parts, ctrl, fold_cfl, s); break;
    case PPK_RIEMANN_HLLC:
      launch_fused<PPK_RIEMANN_HLLC>(h, Uin, Uout, parts, ctrl, fold_cfl, s);
      break;
    default: launch_fused<PPK_RIEMANN_LLF>(h, Uin, Uout, parts, ctrl, fold_cfl, s); break;
  }
  return hipGetLastError();
}

}  // namespace PPK_NS
}  // namespace ppk
