// TEST INFRASTRUCTURE -- CPU restatement of the reference's hydro (Euler) MUSCL-Hancock step,
// implementation version 0, used only as the checker of the library's hydro kernels.  It
// follows the reference's staging: ghost fill, U2 = U, primitives of every cell, fluxes of the
// left faces of every cell of [gw, size-gw] (slopes recomputed per face), then the update.
// Reference: src/muscl/SolverHydroMuscl.cpp:282-405 (godunov_unsplit_impl),
// src/muscl/HydroRunFunctors{2D,3D}.h, src/muscl/HydroBaseFunctor{2D,3D}.h,
// src/shared/RiemannSolvers.h, src/shared/BoundariesFunctors.h:18-420,
// src/muscl/HydroInitFunctors{2D,3D}.h (implode, blast, Kelvin-Helmholtz, Gresho vortex,
// isentropic vortex, four quadrant).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#ifdef _OPENMP
#include <omp.h>
#endif
#include <vector>

#include "../include/ppkmhd_amd.h"

namespace {

// ---- scalar type: plain double, or (HORA_COUNT build) a double that counts its operations --
#ifdef HORA_COUNT
struct OpCounts {
  long long add = 0, mul = 0, div = 0, sqrt = 0, sel = 0;
};
// stages of a step: 0 ghost fill + copy, 1 primitives of every cell, 2 slopes of a cell of the
// flux range, 3 the same recomputed for its left neighbours, 4 Hancock sources (computed in
// each of the 2 * dim traces of a cell), 5 face states + Riemann solve + dt/dx scaling,
// 6 update, 7 CFL maximum
constexpr int kStages = 8;
OpCounts g_stage_ops[kStages];
int g_stage = 0;
#define g_ops (g_stage_ops[g_stage])
struct real {
  double v;
  real() : v(0.0) {}
  real(double x) : v(x) {}
};
inline real operator+(real a, real b) { ++g_ops.add; return real(a.v + b.v); }
inline real operator-(real a, real b) { ++g_ops.add; return real(a.v - b.v); }
inline real operator*(real a, real b) { ++g_ops.mul; return real(a.v * b.v); }
inline real operator/(real a, real b) { ++g_ops.div; return real(a.v / b.v); }
inline real operator-(real a) { return real(-a.v); }
inline real &operator+=(real &a, real b) { a = a + b; return a; }
inline real &operator-=(real &a, real b) { a = a - b; return a; }
inline bool operator<(real a, real b) { return a.v < b.v; }
inline bool operator>(real a, real b) { return a.v > b.v; }
inline bool operator<=(real a, real b) { return a.v <= b.v; }
inline bool operator>=(real a, real b) { return a.v >= b.v; }
inline bool operator!=(real a, real b) { return a.v != b.v; }
inline real SQRT(real a) { ++g_ops.sqrt; return real(std::sqrt(a.v)); }
inline real FABS(real a) { ++g_ops.sel; return real(std::fabs(a.v)); }
inline real FMAX(real a, real b) { ++g_ops.sel; return real(std::fmax(a.v, b.v)); }
inline real FMIN(real a, real b) { ++g_ops.sel; return real(std::fmin(a.v, b.v)); }
inline real COPYSIGN(real a, real b) { return real(std::copysign(a.v, b.v)); }
inline double D(real a) { return a.v; }
#define HORA_STAGE(n) (g_stage = (n))
#else
#define HORA_STAGE(n) ((void)0)
typedef double real;
inline double SQRT(double a) { return std::sqrt(a); }
inline double FABS(double a) { return std::fabs(a); }
inline double FMAX(double a, double b) { return std::fmax(a, b); }
inline double FMIN(double a, double b) { return std::fmin(a, b); }
inline double COPYSIGN(double a, double b) { return std::copysign(a, b); }
inline double D(double a) { return a; }
#endif

enum { ID = 0, IP = 1, IU = 2, IV = 3, IW = 4 };

struct Grid {
  int dim, nv, gw, nx, ny, nz, is, js, ks;
  long long vs;
  explicit Grid(const ppk_mhd_params &p)
      : dim(p.dim), nv(p.dim == 3 ? 5 : 4), gw(p.ghost_width), nx(p.nx), ny(p.ny),
        nz(p.dim == 3 ? p.nz : 1) {
    is = nx + 2 * gw;
    js = ny + 2 * gw;
    ks = dim == 3 ? nz + 2 * gw : 1;
    vs = (long long)is * js * ks;
  }
  long long at(int i, int j, int k, int v) const {
    return (((long long)v * ks + k) * js + j) * (long long)is + i;
  }
};

struct S5 {
  real v[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  real &operator[](int i) { return v[i]; }
  const real &operator[](int i) const { return v[i]; }
};

// ---- boundaries (MakeBoundariesFunctor2D/3D, non-MHD: Dirichlet flips the normal velocity) --
void boundaries(const ppk_mhd_params &p, double *U) {
  const Grid g(p);
  for (int dir = 0; dir < g.dim; ++dir) {
    const int n = dir == 0 ? g.nx : dir == 1 ? g.ny : g.nz;
    for (int side = 0; side < 2; ++side) {
      const int bc = p.bc[2 * dir + side];
      for (int k = 0; k < g.ks; ++k)
        for (int j = 0; j < g.js; ++j)
          for (int i = 0; i < g.is; ++i) {
            const int c = dir == 0 ? i : dir == 1 ? j : k;
            int c0;
            if (side == 0) {
              if (c >= g.gw) continue;
              c0 = bc == PPK_BC_DIRICHLET ? 2 * g.gw - 1 - c : bc == PPK_BC_NEUMANN ? g.gw : n + c;
            } else {
              if (c < n + g.gw) continue;
              c0 = bc == PPK_BC_DIRICHLET ? 2 * n + 2 * g.gw - 1 - c
                   : bc == PPK_BC_NEUMANN ? n + g.gw - 1
                                          : c - n;
            }
            const int i0 = dir == 0 ? c0 : i, j0 = dir == 1 ? c0 : j, k0 = dir == 2 ? c0 : k;
            for (int v = 0; v < g.nv; ++v) {
              const double sign = (bc == PPK_BC_DIRICHLET && v == IU + dir) ? -1.0 : 1.0;
              U[g.at(i, j, k, v)] = U[g.at(i0, j0, k0, v)] * sign;
            }
          }
    }
  }
}

// ---- primitives (computePrimitives) ---------------------------------------------------------
void to_prim(const ppk_mhd_params &p, int nv, const S5 &u, real &c, S5 &q) {
  const real d = FMAX(u[ID], p.smallr);
  const real ux = u[IU] / d, uy = u[IV] / d;
  const real uz = nv == 5 ? u[IW] / d : 0.0;
  const real eken = nv == 5 ? 0.5 * (ux * ux + uy * uy + uz * uz) : 0.5 * (ux * ux + uy * uy);
  const real e = u[IP] / d - eken;
  const real pr = FMAX((p.gamma0 - 1.0) * d * e, d * p.smallp);
  c = SQRT(p.gamma0 * (pr) / d);
  q[ID] = d;
  q[IP] = pr;
  q[IU] = ux;
  q[IV] = uy;
  if (nv == 5) q[IW] = uz;
}

// ---- slopes (slope_unsplit_hydro_2d/3d_scalar) ----------------------------------------------
real minmod_slope(real st, real q, real qp, real qm) {
  const real dlft = st * (q - qm), drgt = st * (qp - q), dcen = 0.5 * (qp - qm);
  const real dsgn = dcen >= 0.0 ? 1.0 : -1.0;
  real dlim = FMIN(FABS(dlft), FABS(drgt));
  if ((dlft * drgt) <= 0.0) dlim = 0.0;
  return dsgn * FMIN(dlim, FABS(dcen));
}

// ---- Riemann solvers (RiemannSolvers.h) -----------------------------------------------------
void flux_of_gdnv(const ppk_mhd_params &p, int nv, const S5 &qg, S5 &f) {
  f[ID] = qg[ID] * qg[IU];
  f[IU] = f[ID] * qg[IU] + qg[IP];
  f[IV] = f[ID] * qg[IV];
  if (nv == 5) f[IW] = f[ID] * qg[IW];
  const real entho = 1.0 / (p.gamma0 - 1.0);
  real ekin = 0.5 * qg[ID] * (qg[IU] * qg[IU] + qg[IV] * qg[IV]);
  if (nv == 5) ekin += 0.5 * qg[ID] * (qg[IW] * qg[IW]);
  const real etot = qg[IP] * entho + ekin;
  f[IP] = qg[IU] * (etot + qg[IP]);
}

void solve_approx(const ppk_mhd_params &p, int nv, const S5 &ql, const S5 &qr, S5 &f) {
  const real g0 = p.gamma0, g6 = p.gamma6;
  const real rl = FMAX(ql[ID], p.smallr), ul = ql[IU], pl = FMAX(ql[IP], rl * p.smallp);
  const real rr = FMAX(qr[ID], p.smallr), ur = qr[IU], pr = FMAX(qr[IP], rr * p.smallp);
  const real cl = g0 * pl * rl, cr = g0 * pr * rr;
  real wl = SQRT(cl), wr = SQRT(cr);
  real pstar = FMAX(((wr * pl + wl * pr) + wl * wr * (ul - ur)) / (wl + wr), 0.0);
  real pold = pstar, conv = 1.0;
  for (int it = 0; it < 10 && conv > 1e-6; ++it) {
    const real wwl = SQRT(cl * (1.0 + g6 * (pold - pl) / pl));
    const real wwr = SQRT(cr * (1.0 + g6 * (pold - pr) / pr));
    const real a = 2.0 * wwl * wwl * wwl / (wwl * wwl + cl);
    const real b = 2.0 * wwr * wwr * wwr / (wwr * wwr + cr);
    const real usl = ul - (pold - pl) / wwl, usr = ur + (pold - pr) / wwr;
    const real delp = FMAX(b * a / (b + a) * (usl - usr), -pold);
    pold = pold + delp;
    conv = FABS(delp / (pold + p.smallpp));
  }
  pstar = pold;
  wl = SQRT(cl * (1.0 + g6 * (pstar - pl) / pl));
  wr = SQRT(cr * (1.0 + g6 * (pstar - pr) / pr));
  const real ustar = 0.5 * (ul + (pl - pstar) / wl + ur - (pr - pstar) / wr);
  const real sgnm = COPYSIGN(1.0, ustar);
  const bool left = sgnm > 0.0;
  const real ro = left ? rl : rr, uo = left ? ul : ur, po = left ? pl : pr, wo = left ? wl : wr;
  const real co = FMAX(p.smallc, SQRT(FABS(g0 * po / ro)));
  const real rstar = FMAX(ro / (1.0 + ro * (po - pstar) / (wo * wo)), p.smallr);
  const real cstar = FMAX(p.smallc, SQRT(FABS(g0 * pstar / rstar)));
  real spout = co - sgnm * uo, spin = cstar - sgnm * ustar;
  const real ushock = wo / ro - sgnm * uo;
  if (pstar >= po) spin = spout = ushock;
  const real scr = FMAX(spout - spin, p.smallc + FABS(spout + spin));
  real frac = 0.5 * (1.0 + (spout + spin) / scr);
  frac = (frac != frac) ? 0.0 : (frac >= 1.0 ? 1.0 : frac <= 0.0 ? 0.0 : frac);
  S5 qg;
  qg[ID] = frac * rstar + (1.0 - frac) * ro;
  qg[IU] = frac * ustar + (1.0 - frac) * uo;
  qg[IP] = frac * pstar + (1.0 - frac) * po;
  if (spout < 0.0) { qg[ID] = ro; qg[IU] = uo; qg[IP] = po; }
  if (spin > 0.0) { qg[ID] = rstar; qg[IU] = ustar; qg[IP] = pstar; }
  qg[IV] = left ? ql[IV] : qr[IV];
  if (nv == 5) qg[IW] = left ? ql[IW] : qr[IW];
  flux_of_gdnv(p, nv, qg, f);
}

void cons_flux(int nv, const S5 &q, real entho, bool hll, S5 &u, S5 &f) {
  u[ID] = q[ID];
  u[IP] = q[IP] * entho + 0.5 * q[ID] * q[IU] * q[IU];
  u[IP] += 0.5 * q[ID] * q[IV] * q[IV];
  if (nv == 5) u[IP] += 0.5 * q[ID] * q[IW] * q[IW];
  u[IU] = q[ID] * q[IU];
  u[IV] = q[ID] * q[IV];
  if (nv == 5) u[IW] = q[ID] * q[IW];
  f[ID] = hll ? u[IU] : u[ID] * q[IU];
  f[IP] = q[IU] * (u[IP] + q[IP]);
  f[IU] = q[IP] + u[IU] * q[IU];
  f[IV] = f[ID] * q[IV];
  if (nv == 5) f[IW] = f[ID] * q[IW];
}

void solve_llf_hll(const ppk_mhd_params &p, int nv, const S5 &ql, const S5 &qr, bool hll,
                   S5 &f) {
  const real g0 = p.gamma0, entho = 1.0 / (g0 - 1.0);
  const real rl = FMAX(ql[ID], p.smallr), ul = ql[IU], pl = FMAX(ql[IP], rl * p.smallp);
  const real rr = FMAX(qr[ID], p.smallr), ur = qr[IU], pr = FMAX(qr[IP], rr * p.smallp);
  const real cl = SQRT(g0 * pl / rl), cr = SQRT(g0 * pr / rr);
  S5 uL, uR, fL, fR;
  cons_flux(nv, ql, entho, hll, uL, fL);
  cons_flux(nv, qr, entho, hll, uR, fR);
  if (!hll) {
    const real cmax = FMAX(FABS(ul) + cl, FABS(ur) + cr);
    for (int n = 0; n < nv; ++n) f[n] = 0.5 * (fL[n] + fR[n] - cmax * (uR[n] - uL[n]));
  } else {
    const real SL = FMIN(FMIN(ul, ur) - FMAX(cl, cr), 0.0);
    const real SR = FMAX(FMAX(ul, ur) + FMAX(cl, cr), 0.0);
    for (int n = 0; n < nv; ++n)
      f[n] = (SR * fL[n] - SL * fR[n] + SR * SL * (uR[n] - uL[n])) / (SR - SL);
  }
}

void solve_hllc(const ppk_mhd_params &p, int nv, const S5 &ql, const S5 &qr, S5 &f) {
  const real g0 = p.gamma0, entho = 1.0 / (g0 - 1.0);
  const real rl = FMAX(ql[ID], p.smallr), pl = FMAX(ql[IP], rl * p.smallp), ul = ql[IU];
  real ecinl = 0.5 * rl * ul * ul;
  ecinl += 0.5 * rl * ql[IV] * ql[IV];
  if (nv == 5) ecinl += 0.5 * rl * ql[IW] * ql[IW];
  const real etotl = pl * entho + ecinl, ptotl = pl;
  const real rr = FMAX(qr[ID], p.smallr), pr = FMAX(qr[IP], rr * p.smallp), ur = qr[IU];
  real ecinr = 0.5 * rr * ur * ur;
  ecinr += 0.5 * rr * qr[IV] * qr[IV];
  if (nv == 5) ecinr += 0.5 * rr * qr[IW] * qr[IW];
  const real etotr = pr * entho + ecinr, ptotr = pr;
  const real cfl_ = SQRT(FMAX(g0 * pl / rl, p.smallc * p.smallc));
  const real cfr = SQRT(FMAX(g0 * pr / rr, p.smallc * p.smallc));
  const real SL = FMIN(ul, ur) - FMAX(cfl_, cfr);
  const real SR = FMAX(ul, ur) + FMAX(cfl_, cfr);
  const real rcl = rl * (ul - SL), rcr = rr * (SR - ur);
  const real ustar = (rcr * ur + rcl * ul + (ptotl - ptotr)) / (rcr + rcl);
  const real ptotstar = (rcr * ptotl + rcl * ptotr + rcl * rcr * (ul - ur)) / (rcr + rcl);
  const real rstarl = rl * (SL - ul) / (SL - ustar);
  const real etotstarl = ((SL - ul) * etotl - ptotl * ul + ptotstar * ustar) / (SL - ustar);
  const real rstarr = rr * (SR - ur) / (SR - ustar);
  const real etotstarr = ((SR - ur) * etotr - ptotr * ur + ptotstar * ustar) / (SR - ustar);
  real ro, uo, ptoto, etoto;
  if (SL > 0.0) { ro = rl; uo = ul; ptoto = ptotl; etoto = etotl; }
  else if (ustar > 0.0) { ro = rstarl; uo = ustar; ptoto = ptotstar; etoto = etotstarl; }
  else if (SR > 0.0) { ro = rstarr; uo = ustar; ptoto = ptotstar; etoto = etotstarr; }
  else { ro = rr; uo = ur; ptoto = ptotr; etoto = etotr; }
  f[ID] = ro * uo;
  f[IU] = ro * uo * uo + ptoto;
  f[IP] = (etoto + ptoto) * uo;
  f[IV] = f[ID] > 0.0 ? f[ID] * ql[IV] : f[ID] * qr[IV];
  if (nv == 5) f[IW] = f[ID] > 0.0 ? f[ID] * ql[IW] : f[ID] * qr[IW];
}

void solve(const ppk_mhd_params &p, int nv, const S5 &ql, const S5 &qr, S5 &f) {
  if (p.riemann == PPK_RIEMANN_APPROX) solve_approx(p, nv, ql, qr, f);
  else if (p.riemann == PPK_RIEMANN_HLL) solve_llf_hll(p, nv, ql, qr, true, f);
  else if (p.riemann == PPK_RIEMANN_HLLC) solve_hllc(p, nv, ql, qr, f);
  else solve_llf_hll(p, nv, ql, qr, false, f);
}

// ---- the step -------------------------------------------------------------------------------
struct HydroSolver {
  ppk_mhd_params p;
  Grid g;
  std::vector<double> U, U2, Q, F[3];
  double t = 0.0, dt = 0.0;
  long long it = 0;
  // gravity ([gravity] static or point: SolverBase.cpp:138-141) and its per-cell field
  // (VectorField gravity, dim components): the one the problem's initial condition writes, 0
  // for problems that write none (the reference's view is zero-initialised)
  std::vector<double> G;
  bool gravity() const { return p.gravity_static != 0; }
  real gf(int i, int j, int k, int d) const { return G[(long long)d * g.vs + g.at(i, j, k, 0)]; }
  explicit HydroSolver(const ppk_mhd_params &pp) : p(pp), g(pp) {
    U.assign(g.nv * g.vs, 0.0);
    G.assign(g.dim * g.vs, 0.0);
    U2 = U;
    Q = U;
    for (int d = 0; d < 3; ++d) F[d] = U;
  }
  S5 cell(const std::vector<double> &A, int i, int j, int k) const {
    S5 s;
    for (int v = 0; v < g.nv; ++v) s[v] = A[g.at(i, j, k, v)];
    return s;
  }
  // compute_dt_local (SolverHydroMuscl.h:650-690): interior cells of the current state; with
  // gravity ComputeDtGravityFunctor2D/3D (HydroRunFunctors2D.h:120-170, 3D.h:130-185): the
  // summed speeds scaled by k / (sqrt(1 + 2k) - 1), k = |g|_1 cfl dx / u^2 floored at 1e-4,
  // over the smallest cell size
  real inv_dt(const std::vector<double> &A) const {
    real inv = 0.0;
    if (gravity()) {
      real dxm = FMIN(real(p.dx), real(p.dy));
      if (g.dim == 3) dxm = FMIN(dxm, real(p.dz));
      for (int k = g.dim == 3 ? g.gw : 0; k < (g.dim == 3 ? g.ks - g.gw : 1); ++k)
        for (int j = g.gw; j < g.js - g.gw; ++j)
          for (int i = g.gw; i < g.is - g.gw; ++i) {
            S5 q;
            real c = 0.0;
            to_prim(p, g.nv, cell(A, i, j, k), c, q);
            real velocity = 0.0;
            velocity += c + FABS(q[IU]);
            velocity += c + FABS(q[IV]);
            if (g.dim == 3) velocity += c + FABS(q[IW]);
            real kk = g.dim == 3
                          ? FABS(gf(i, j, k, 0)) + FABS(gf(i, j, k, 1)) + FABS(gf(i, j, k, 2))
                          : FABS(gf(i, j, k, 0)) + FABS(gf(i, j, k, 1));
            kk = kk * (p.cfl * dxm / (velocity * velocity));
            kk = FMAX(kk, real(1e-4));
            velocity = velocity * (kk / (SQRT(1.0 + 2.0 * kk) - 1.0));
            inv = FMAX(inv, velocity / dxm);
          }
      return inv;
    }
    for (int k = g.dim == 3 ? g.gw : 0; k < (g.dim == 3 ? g.ks - g.gw : 1); ++k)
      for (int j = g.gw; j < g.js - g.gw; ++j)
        for (int i = g.gw; i < g.is - g.gw; ++i) {
          S5 q;
          real c = 0.0;
          to_prim(p, g.nv, cell(A, i, j, k), c, q);
          const real vx = c + FABS(q[IU]), vy = c + FABS(q[IV]);
          if (g.dim == 3) {
            const real vz = c + FABS(q[IW]);
            inv = FMAX(inv, vx / p.dx + vy / p.dy + vz / p.dz);
          } else {
            inv = FMAX(inv, vx / p.dx + vy / p.dy);
          }
        }
    return inv;
  }
  // slopes of cell (i,j,k) of Q
  void cell_slopes(int i, int j, int k, S5 &dx, S5 &dy, S5 &dz) const {
    const S5 q = cell(Q, i, j, k);
    if (p.slope_type == 0) { dx = dy = dz = S5(); return; }
    const S5 qpx = cell(Q, i + 1, j, k), qmx = cell(Q, i - 1, j, k);
    const S5 qpy = cell(Q, i, j + 1, k), qmy = cell(Q, i, j - 1, k);
    for (int v = 0; v < g.nv; ++v) {
      dx[v] = minmod_slope(p.slope_type, q[v], qpx[v], qmx[v]);
      dy[v] = minmod_slope(p.slope_type, q[v], qpy[v], qmy[v]);
    }
    if (g.dim == 3) {
      const S5 qpz = cell(Q, i, j, k + 1), qmz = cell(Q, i, j, k - 1);
      for (int v = 0; v < g.nv; ++v) dz[v] = minmod_slope(p.slope_type, q[v], qpz[v], qmz[v]);
    }
  }
  // trace_unsplit_{2d,3d}_along_dir of a cell to face (dir, side: 0 min, 1 max)
  S5 trace(const S5 &q, const S5 &dx, const S5 &dy, const S5 &dz, int dir, int side,
           real dtdx, real dtdy, real dtdz) const {
    HORA_STAGE(4);
    const real g0 = p.gamma0;
    const real r = q[ID], pr = q[IP], u = q[IU], v = q[IV], w = g.dim == 3 ? q[IW] : 0.0;
    real sr0, sp0, su0, sv0, sw0 = 0.0;
    if (g.dim == 3) {
      sr0 = -u * dx[ID] - v * dy[ID] - w * dz[ID] - (dx[IU] + dy[IV] + dz[IW]) * r;
      sp0 = -u * dx[IP] - v * dy[IP] - w * dz[IP] - (dx[IU] + dy[IV] + dz[IW]) * g0 * pr;
      su0 = -u * dx[IU] - v * dy[IU] - w * dz[IU] - (dx[IP]) / r;
      sv0 = -u * dx[IV] - v * dy[IV] - w * dz[IV] - (dy[IP]) / r;
      sw0 = -u * dx[IW] - v * dy[IW] - w * dz[IW] - (dz[IP]) / r;
    } else {
      sr0 = -u * dx[ID] - v * dy[ID] - (dx[IU] + dy[IV]) * r;
      sp0 = -u * dx[IP] - v * dy[IP] - (dx[IU] + dy[IV]) * g0 * pr;
      su0 = -u * dx[IU] - v * dy[IU] - (dx[IP]) / r;
      sv0 = -u * dx[IV] - v * dy[IV] - (dy[IP]) / r;
    }
    HORA_STAGE(5);
    const S5 &d = dir == 0 ? dx : dir == 1 ? dy : dz;
    const real dtd = dir == 0 ? dtdx : dir == 1 ? dtdy : dtdz;
    const real s0[5] = {sr0, sp0, su0, sv0, sw0};
    const real c0[5] = {r, pr, u, v, w};
    S5 out;
    for (int n = 0; n < g.nv; ++n)
      out[n] = side == 0 ? c0[n] - 0.5 * d[n] + s0[n] * dtd * 0.5
                         : c0[n] + 0.5 * d[n] + s0[n] * dtd * 0.5;
    out[ID] = FMAX(p.smallr, out[ID]);
    return out;
  }
  // godunov_unsplit_impl, implementation version 0 (SolverHydroMuscl.cpp:282-405)
  void godunov(std::vector<double> &in, std::vector<double> &out, double dt_) {
    HORA_STAGE(0);
    boundaries(p, in.data());
    out = in;
    HORA_STAGE(1);
    for (int k = 0; k < g.ks; ++k)
      for (int j = 0; j < g.js; ++j)
        for (int i = 0; i < g.is; ++i) {
          S5 q;
          real c;
          to_prim(p, g.nv, cell(in, i, j, k), c, q);
          for (int v = 0; v < g.nv; ++v) Q[g.at(i, j, k, v)] = D(q[v]);
        }
    const real dtdx = dt_ / p.dx, dtdy = dt_ / p.dy, dtdz = g.dim == 3 ? dt_ / p.dz : 0.0;  // once a step
    const real dtd[3] = {dtdx, dtdy, dtdz};
    const int k0 = g.dim == 3 ? g.gw : 0, k1 = g.dim == 3 ? g.ks - g.gw : 0;
    // ComputeAndStoreFluxesFunctor: the left face of every cell of [gw, size-gw]
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = k0; k <= k1; ++k)
      for (int j = g.gw; j <= g.js - g.gw; ++j)
        for (int i = g.gw; i <= g.is - g.gw; ++i) {
          S5 dx, dy, dz;
          HORA_STAGE(2);
          cell_slopes(i, j, k, dx, dy, dz);
          const S5 qloc = cell(Q, i, j, k);
          for (int dir = 0; dir < g.dim; ++dir) {
            const int ni = i - (dir == 0), nj = j - (dir == 1), nk = k - (dir == 2);
            S5 ndx, ndy, ndz;
            HORA_STAGE(3);
            cell_slopes(ni, nj, nk, ndx, ndy, ndz);
            S5 qright = trace(qloc, dx, dy, dz, dir, 0, dtdx, dtdy, dtdz);
            S5 qleft = trace(cell(Q, ni, nj, nk), ndx, ndy, ndz, dir, 1, dtdx, dtdy, dtdz);
            if (gravity()) {  // gravity predictor, half a time step (HydroRunFunctors2D.h:650-660):
                              // each state with its own cell's field
              for (int d = 0; d < g.dim; ++d) {
                qleft[IU + d] += 0.5 * dt_ * gf(ni, nj, nk, d);
                qright[IU + d] += 0.5 * dt_ * gf(i, j, k, d);
              }
            }
            const int nrm = IU + dir;  // rotate the normal velocity into IU
            std::swap(qleft[IU], qleft[nrm]);
            std::swap(qright[IU], qright[nrm]);
            S5 f;
            solve(p, g.nv, qleft, qright, f);
            std::swap(f[IU], f[nrm]);  // and back
            for (int v = 0; v < g.nv; ++v) F[dir][g.at(i, j, k, v)] = D(f[v] * dtd[dir]);
          }
        }
    // UpdateFunctor: +F(i) - F(i+1) per direction, in that order, variable by variable
    HORA_STAGE(6);
    for (int k = k0; k < (g.dim == 3 ? k1 : 1); ++k)
      for (int j = g.gw; j < g.js - g.gw; ++j)
        for (int i = g.gw; i < g.is - g.gw; ++i)
          for (int v = 0; v < g.nv; ++v) {
            real o = out[g.at(i, j, k, v)];
            o += F[0][g.at(i, j, k, v)];
            o -= F[0][g.at(i + 1, j, k, v)];
            o += F[1][g.at(i, j, k, v)];
            o -= F[1][g.at(i, j + 1, k, v)];
            if (g.dim == 3) {
              o += F[2][g.at(i, j, k, v)];
              o -= F[2][g.at(i, j, k + 1, v)];
            }
            out[g.at(i, j, k, v)] = D(o);
          }
    // GravitySourceTermFunctor2D/3D (HydroRunFunctors2D.h:1299-1331, 3D.h:1463-1500): momentum
    // by dt g times the mean of the old and new densities, energy by the change of kinetic
    // energy; the 2D functor floors the new density at smallr, the 3D one does not
    if (gravity()) {
      for (int k = k0; k < (g.dim == 3 ? k1 : 1); ++k)
        for (int j = g.gw; j < g.js - g.gw; ++j)
          for (int i = g.gw; i < g.is - g.gw; ++i) {
            const real rhoOld = in[g.at(i, j, k, ID)];
            const real rhoNew = g.dim == 3 ? real(out[g.at(i, j, k, ID)])
                                           : FMAX(real(p.smallr), real(out[g.at(i, j, k, ID)]));
            real m[3] = {out[g.at(i, j, k, IU)], out[g.at(i, j, k, IV)],
                         g.dim == 3 ? out[g.at(i, j, k, IW)] : 0.0};
            const real ekin_old = g.dim == 3
                                      ? 0.5 * (m[0] * m[0] + m[1] * m[1] + m[2] * m[2]) / rhoNew
                                      : 0.5 * (m[0] * m[0] + m[1] * m[1]) / rhoNew;
            for (int d = 0; d < g.dim; ++d) m[d] += 0.5 * dt_ * gf(i, j, k, d) * (rhoOld + rhoNew);
            for (int d = 0; d < g.dim; ++d) out[g.at(i, j, k, IU + d)] = D(m[d]);
            const real ekin_new = g.dim == 3
                                      ? 0.5 * (m[0] * m[0] + m[1] * m[1] + m[2] * m[2]) / rhoNew
                                      : 0.5 * (m[0] * m[0] + m[1] * m[1]) / rhoNew;
            real e = out[g.at(i, j, k, IP)];
            e += (ekin_new - ekin_old);
            out[g.at(i, j, k, IP)] = D(e);
          }
    }
  }
  // SolverBase::next_iteration (src/shared/SolverBase.cpp:161-219) with the hydro dt
  void step(double dt_fixed, int dt_mode) {
    std::vector<double> &cur = (it % 2 == 0) ? U : U2;
    std::vector<double> &nxt = (it % 2 == 0) ? U2 : U;
    if (dt_mode == PPK_DT_CFL) {
      HORA_STAGE(7);
      dt = p.cfl / D(inv_dt(cur));
      if (t + dt > p.tEnd) dt = p.tEnd - t;
    } else {  // the library clamps a fixed dt to tEnd as the CFL one
      dt = dt_fixed;
      if (t + dt > p.tEnd) dt = p.tEnd - t;
    }
    godunov(cur, nxt, dt);
    ++it;
    t += dt;
  }
};

// ---- initial conditions (HydroInitFunctors2D.h:35-660, HydroInitFunctors3D.h:66-530) --------
// four-quadrant states, (density, x velocity, y velocity, pressure) of the quadrants U0..U3
// of src/shared/problems/initRiemannConfig2d.h:42-514 (U0 upper right, then counter-clockwise)
struct Quad { double d, u, v, p; };
const Quad kFourQuadrant[19][4] = {
    {{1.0, 0.0, 0.0, 1.0}, {0.5197, -0.7259, 0.0, 0.4},
     {0.1072, -0.7259, -1.4045, 0.0439}, {0.2579, 0.0, -1.4045, 0.15}},  // configuration 1
    {{1.0, 0.0, 0.0, 1.0}, {0.5197, -0.7259, 0.0, 0.4},
     {1.0, -0.7259, -0.7259, 1.0}, {0.5197, 0.0, -0.7259, 0.4}},  // configuration 2
    {{1.5, 0.0, 0.0, 1.5}, {0.5323, 1.206, 0.0, 0.3},
     {0.138, 1.206, 1.206, 0.029}, {0.5323, 0.0, 1.206, 0.3}},  // configuration 3
    {{1.1, 0.0, 0.0, 1.1}, {0.5065, 0.8939, 0.0, 0.35},
     {1.1, 0.8939, 0.8939, 1.1}, {0.5065, 0.0, 0.8939, 0.35}},  // configuration 4
    {{1.0, -0.75, -0.5, 1.0}, {2.0, -0.75, 0.5, 1.0},
     {1.0, 0.75, 0.5, 1.0}, {3.0, 0.75, -0.5, 1.0}},  // configuration 5
    {{1.0, 0.75, -0.5, 1.0}, {2.0, 0.75, 0.5, 0.5},
     {1.0, -0.75, 0.5, 1.0}, {3.0, -0.75, -0.5, 1.0}},  // configuration 6
    {{1.0, 0.1, 0.1, 1.0}, {0.5197, -0.6259, 0.1, 0.4},
     {0.8, 0.1, 0.1, 0.4}, {0.5197, 0.1, -0.6259, 0.4}},  // configuration 7
    {{0.5197, 0.1, 0.1, 0.4}, {1.0, -0.6259, 0.1, 1.0},
     {0.8, 0.1, 0.1, 1.0}, {1.0, 0.1, -0.6259, 1.0}},  // configuration 8
    {{1.0, 0.0, 0.3, 1.0}, {2.0, 0.0, -0.3, 1.0},
     {1.039, 0.0, -0.8133, 0.4}, {0.5197, 0.0, -0.4259, 0.4}},  // configuration 9
    {{1.0, 0.0, 0.4297, 1.0}, {0.5, 0.0, 0.6076, 1.0},
     {0.2281, 0.0, -0.6076, 0.3333}, {0.4562, 0.0, -0.4259, 0.3333}},  // configuration 10
    {{1.0, 0.1, 0.0, 1.0}, {0.5313, 0.8276, 0.0, 0.4},
     {0.8, 0.1, 0.0, 0.4}, {0.5313, 0.1, 0.7276, 0.4}},  // configuration 11
    {{0.5313, 0.0, 0.0, 0.4}, {1.0, 0.7276, 0.0, 1.0},
     {0.8, 0.0, 0.0, 1.0}, {1.0, 0.0, 0.7276, 1.0}},  // configuration 12
    {{1.0, 0.0, -0.3, 1.0}, {2.0, 0.0, 0.3, 1.0},
     {1.0625, 0.0, 0.8145, 0.4}, {0.5313, 0.0, 0.4276, 0.4}},  // configuration 13
    {{2.0, 0.0, -0.5606, 8.0}, {1.0, 0.0, -1.2172, 8.0},
     {0.4736, 0.0, 1.2172, 2.6667}, {0.9474, 0.0, 1.1606, 2.6667}},  // configuration 14
    {{1.0, 0.1, -0.3, 1.0}, {0.5197, -0.6259, -0.3, 0.4},
     {0.8, 0.1, -0.3, 0.4}, {0.5313, 0.1, 0.4276, 0.4}},  // configuration 15
    {{0.5313, 0.1, 0.1, 0.4}, {1.0222, -0.6179, 0.1, 1.0},
     {0.8, 0.1, 0.1, 1.0}, {1.0, 0.1, 0.8276, 1.0}},  // configuration 16
    {{1.0, 0.0, -0.4, 1.0}, {2.0, 0.0, -0.3, 1.0},
     {1.0625, 0.0, 0.2145, 0.4}, {0.5197, 0.0, -1.1259, 0.4}},  // configuration 17
    {{1.0, 0.0, 1.0, 1.0}, {2.0, 0.0, -0.3, 1.0},
     {1.0625, 0.0, 0.2145, 0.4}, {0.5197, 0.0, 0.2741, 0.4}},  // configuration 18
    {{1.0, 0.0, 0.3, 1.0}, {2.0, 0.0, -0.3, 1.0},
     {1.0625, 0.0, 0.2145, 0.4}, {0.5197, 0.0, -0.4259, 0.4}},  // configuration 19
};
// getRiemannConfig2d's switch falls from its last case (18) into the default (configuration
// 1): only numbers 0..17 select their own row
const Quad *four_quadrant_config(int n) { return kFourQuadrant[(n >= 0 && n < 18) ? n : 0]; }

// counter generator of the random Kelvin-Helmholtz layer (splitmix64 of seed, padded-box cell,
// draw), the library's documented stand-in for the reference's thread-mapped XorShift pool
double kh_rand(uint64_t seed, uint64_t cell, uint64_t draw) {
  uint64_t x = seed * 0x9E3779B97F4A7C15ull + cell * 0xD1B54A32D192ED03ull + draw * 0x8CB92BA72F3D8DD7ull;
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  x = x ^ (x >> 31);
  return (double)(x >> 11) * (1.0 / 9007199254740992.0);
}

// the gravity field the problems' initial conditions write: uniform for Rayleigh-Taylor and
// the rising bubble, PointSourceGravity::eval (src/shared/problems/PointSourceGravity.h)
// at disk-relative coordinates for the disk (InitDiskFunctor2D/3D call it with x - xc ...),
// zero for the others
void gravity_field(const ppk_mhd_params &p, const ppk_mhd_problem &pr, double *G) {
  const Grid g(p);
  for (int k = 0; k < g.ks; ++k)
    for (int j = 0; j < g.js; ++j)
      for (int i = 0; i < g.is; ++i) {
        double v[3] = {0, 0, 0};
        if (pr.kind == PPK_PROBLEM_RAYLEIGH_TAYLOR) {
          v[0] = pr.rti_gx; v[1] = pr.rti_gy; v[2] = pr.rti_gz;
        } else if (pr.kind == PPK_PROBLEM_RISING_BUBBLE) {
          v[0] = pr.rb_gx; v[1] = pr.rb_gy; v[2] = pr.rb_gz;
        } else if (pr.kind == PPK_PROBLEM_DISK) {
          const double x = p.xmin + p.dx / 2 + (i - g.gw) * p.dx - pr.disk_xc;
          const double y = p.ymin + p.dy / 2 + (j - g.gw) * p.dy - pr.disk_yc;
          const double z = g.dim == 3 ? p.zmin + p.dz / 2 + (k - g.gw) * p.dz - pr.disk_zc : 0.0;
          const double dx = x - pr.pg_x, dy = y - pr.pg_y, dz = z - pr.pg_z;
          const double r = std::sqrt(dx * dx + dy * dy + dz * dz + pr.pg_eps * pr.pg_eps);
          const double GM2 = pr.pg_gm / (r * r * r);
          v[0] = GM2 * (pr.pg_x - x);
          v[1] = GM2 * (pr.pg_y - y);
          v[2] = GM2 * (pr.pg_z - z);
        }
        for (int d = 0; d < g.dim; ++d) G[(long long)d * g.vs + g.at(i, j, k, 0)] = v[d];
      }
}

void init(const ppk_mhd_params &p, const ppk_mhd_problem &pr, double *U) {
  const Grid g(p);
  std::fill(U, U + g.nv * g.vs, 0.0);
  for (int k = 0; k < g.ks; ++k)
    for (int j = 0; j < g.js; ++j)
      for (int i = 0; i < g.is; ++i) {
        const double x = p.xmin + p.dx / 2 + (i - g.gw) * p.dx;
        const double y = p.ymin + p.dy / 2 + (j - g.gw) * p.dy;
        const double z = g.dim == 3 ? p.zmin + p.dz / 2 + (k - g.gw) * p.dz : 0.0;
        const double PI = 3.14159265358979323846;  // M_PI
        double d, e, u, v, w;
        if (pr.kind == PPK_PROBLEM_KELVIN_HELMHOLTZ || pr.kind == PPK_PROBLEM_GRESHO_VORTEX ||
            pr.kind == PPK_PROBLEM_ISENTROPIC_VORTEX || pr.kind == PPK_PROBLEM_FOUR_QUADRANT ||
            pr.kind == PPK_PROBLEM_RAYLEIGH_TAYLOR || pr.kind == PPK_PROBLEM_RISING_BUBBLE ||
            pr.kind == PPK_PROBLEM_DISK) {
          // these store momenta and total energy directly
          double st[5] = {0, 0, 0, 0, 0};
          if (pr.kind == PPK_PROBLEM_KELVIN_HELMHOLTZ) {  // InitKelvinHelmholtzFunctor2D/3D
            const double pres = pr.kh_pressure, gm1 = p.gamma0 - 1.0;
            if (pr.kh_p_rand) {
              const double sn = g.dim == 2 ? (y - p.ymin) / (p.ymax - p.ymin)
                                           : (z - p.zmin) / (p.zmax - p.zmin);
              double dd, uu, vv = 0.0, ww = 0.0;
              if (sn < 0.25 || sn > 0.75) { dd = pr.kh_d_out; uu = pr.kh_vflow_out; }
              else { dd = pr.kh_d_in; uu = pr.kh_vflow_in; }
              const uint64_t cell = ((uint64_t)k * g.js + j) * g.is + i;
              uu += pr.kh_amplitude * (kh_rand(pr.kh_seed, cell, 0) - 0.5);
              vv += pr.kh_amplitude * (kh_rand(pr.kh_seed, cell, 1) - 0.5);
              if (g.dim == 3) ww += pr.kh_amplitude * (kh_rand(pr.kh_seed, cell, 2) - 0.5);
              st[ID] = dd;
              st[IU] = dd * uu;
              st[IV] = dd * vv;
              st[IW] = dd * ww;
              st[IP] = g.dim == 2 ? pres / gm1 + 0.5 * dd * (uu * uu + vv * vv)
                                  : pres / gm1 + 0.5 * dd * (uu * uu + vv * vv + ww * ww);
            } else if (pr.kh_p_sine_rob) {
              const double c = g.dim == 2 ? y : z;
              const double ramp = 1.0 / (1.0 + std::exp(2 * (c - 0.25) / pr.kh_delta)) +
                                  1.0 / (1.0 + std::exp(2 * (0.75 - c) / pr.kh_delta));
              const double dd = pr.kh_d_in + ramp * (pr.kh_d_out - pr.kh_d_in);
              const double uu = pr.kh_vflow_in + ramp * (pr.kh_vflow_out - pr.kh_vflow_in);
              const int n = pr.kh_mode;
              if (g.dim == 2) {
                const double vv = pr.kh_w0 * std::sin(n * PI * x);
                st[ID] = dd; st[IU] = dd * uu; st[IV] = dd * vv;
                st[IP] = pres / gm1 + 0.5 * dd * (uu * uu + vv * vv);
              } else {
                const double vv = pr.kh_vflow_in / 2 + ramp * (pr.kh_vflow_out / 2 - pr.kh_vflow_in / 2);
                const double ww = pr.kh_w0 * std::sin(n * PI * x) * std::sin(n * PI * y);
                st[ID] = dd; st[IU] = dd * uu; st[IV] = dd * vv; st[IW] = dd * ww;
                st[IP] = pres / gm1 + 0.5 * dd * (uu * uu + vv * vv + ww * ww);
              }
            }
          } else if (pr.kind == PPK_PROBLEM_GRESHO_VORTEX) {  // InitGreshoVortexFunctor2D/3D
            const double rho0 = pr.gresho_rho0;
            const double p0 = rho0 / (p.gamma0 * pr.gresho_ma * pr.gresho_ma);
            const double r = std::sqrt(x * x + y * y), th = std::atan2(y, x);
            double uphi, pp;
            if (r < 0.2) { uphi = 5 * r; pp = p0 + 25 / 2.0 * r * r; }
            else if (r < 0.4) {
              uphi = 2 - 5 * r;
              pp = p0 + 25 / 2.0 * r * r + 4 * (1 - 5 * r - std::log(0.2) + std::log(r));
            } else { uphi = 0; pp = p0 - 2 + 4 * std::log(2.0); }
            st[ID] = rho0;
            st[IU] = rho0 * (-std::sin(th) * uphi);
            st[IV] = rho0 * (std::cos(th) * uphi);
            if (g.dim == 3) st[IW] = rho0 * (std::cos(th) * uphi);  // the reference's 3D choice
            double m2 = st[IU] * st[IU] + st[IV] * st[IV];
            if (g.dim == 3) m2 = m2 + st[IW] * st[IW];
            st[IP] = pp / (p.gamma0 - 1.0) + 0.5 * m2 / st[ID];
          } else if (pr.kind == PPK_PROBLEM_RAYLEIGH_TAYLOR) {
            // RayleighTaylorInstabilityFunctor2D/3D (HydroInitFunctors2D.h:695-780,
            // 3D.h:578-676): heavy fluid above the mid-plane, a cosine bump in the vertical
            // velocity (stored in the momentum slot as the reference does), hydrostatic
            // pressure P0 + rho g.x with P0 = 1
            const double Lx = p.xmax - p.xmin, Ly = p.ymax - p.ymin, P0 = 1.0;
            if (g.dim == 2) {
              st[ID] = y > (p.ymin + p.ymax) / 2 ? pr.rti_d1 : pr.rti_d0;
              st[IU] = 0.0;
              st[IV] = pr.rti_amplitude * (1 + std::cos(2 * PI * x / Lx)) *
                       (1 + std::cos(2 * PI * y / Ly)) / 4;
              st[IP] = (P0 + st[ID] * (pr.rti_gx * x + pr.rti_gy * y)) / (p.gamma0 - 1.0);
            } else {
              st[ID] = z > (p.zmin + p.zmax) / 2 ? pr.rti_d1 : pr.rti_d0;
              st[IU] = 0.0;
              st[IV] = 0.0;
              st[IW] = pr.rti_amplitude * (1 + std::cos(2 * PI * x / Lx)) *
                       (1 + std::cos(2 * PI * y / Ly)) / 4;
              st[IP] = (P0 + st[ID] * (pr.rti_gx * x + pr.rti_gy * y + pr.rti_gz * z)) /
                       (p.gamma0 - 1);
            }
          } else if (pr.kind == PPK_PROBLEM_DISK) {
            // InitDiskFunctor2D/3D (HydroInitFunctors2D.h:907-1030, 3D.h:807-935) with
            // DiskParams::radial_speed_of_sound
            const double xr = x - pr.disk_xc, yr = y - pr.disk_yc;
            const double zr = g.dim == 3 ? z - pr.disk_zc : 0.0;
            const double GM = pr.pg_gm, eps = pr.pg_eps;
            const double r2 = xr * xr + yr * yr;
            const double r = std::sqrt(r2);
            const double r_soft = std::sqrt(r2 + eps * eps);
            const double r_sph = std::sqrt(r2 + eps * eps + zr * zr);
            const double x_soft = xr * (r_soft / r);
            const double y_soft = yr * (r_soft / r);
            double csound;
            if (r < pr.disk_radius_inner)
              csound = pr.disk_ref_sound_speed / std::sqrt(pr.disk_radius_inner / pr.disk_radius);
            else
              csound = pr.disk_ref_sound_speed / std::sqrt(r / pr.disk_radius);
            double omega2;
            if (r >= pr.disk_radius_inner)
              omega2 = GM / (r_soft * r_soft * r_sph) -
                       3.5 * (csound / r_soft) * (csound / r_soft);
            else
              omega2 = GM / (r_soft * r_soft * r_soft) -
                       2.5 * (csound / r_soft) * (csound / r_soft);
            const double omega = std::sqrt(std::fmax(omega2, 0.0));
            double rho = pr.disk_ref_density * std::pow(r_soft / pr.disk_radius, -2.5) *
                         std::exp(GM / (csound * csound) * (1. / r_sph - 1. / r_soft));
            if (r > pr.disk_radius || std::fabs(zr) > 0.5 * pr.disk_radius) {
              if (r > pr.disk_radius + pr.disk_contrast_width ||
                  std::fabs(zr) > 0.5 * pr.disk_radius + pr.disk_contrast_width) {
                rho /= pr.disk_contrast_density;
              } else {
                const double cutdown =
                    std::fmax((r - pr.disk_radius) / pr.disk_contrast_width,
                              (std::fabs(zr) - 0.5 * pr.disk_radius) / pr.disk_contrast_width);
                rho *= std::pow(pr.disk_contrast_density, -cutdown);
              }
            }
            rho = std::fmax(rho, 1e-6);
            st[ID] = rho;
            st[IU] = -y_soft * omega * rho;
            st[IV] = x_soft * omega * rho;
            const double pressure = csound * csound * rho;
            st[IP] = pressure / (p.gamma0 - 1.0) + 0.5 * rho * omega2 * (r2 + eps * eps);
          } else if (pr.kind == PPK_PROBLEM_RISING_BUBBLE) {
            // RisingBubbleFunctor2D/3D (HydroInitFunctors2D.h:840-900, 3D.h:715-800): a
            // sphere of density din in dout at rest, hydrostatic pressure P0 + rho g.x
            const double P0 = 1.0;
            const double r = g.dim == 2
                                 ? std::sqrt((x - pr.rb_x0) * (x - pr.rb_x0) +
                                             (y - pr.rb_y0) * (y - pr.rb_y0))
                                 : std::sqrt((x - pr.rb_x0) * (x - pr.rb_x0) +
                                             (y - pr.rb_y0) * (y - pr.rb_y0) +
                                             (z - pr.rb_z0) * (z - pr.rb_z0));
            st[ID] = r < pr.rb_R ? pr.rb_din : pr.rb_dout;
            st[IU] = st[IV] = st[IW] = 0.0;
            st[IP] = g.dim == 2
                         ? (P0 + st[ID] * (pr.rb_gx * x + pr.rb_gy * y)) / (p.gamma0 - 1.0)
                         : (P0 + st[ID] * (pr.rb_gx * x + pr.rb_gy * y + pr.rb_gz * z)) /
                               (p.gamma0 - 1.0);
          } else if (pr.kind == PPK_PROBLEM_FOUR_QUADRANT) {
            // InitFourQuadrantFunctor2D (HydroInitFunctors2D.h:496-559) of states converted by
            // primToCons_2D (initRiemannConfig2d.h:25-38)
            const Quad *cf = four_quadrant_config(pr.fq_config);
            Quad q;
            if (x < pr.fq_x) q = y < pr.fq_y ? cf[2] : cf[1];
            else q = y < pr.fq_y ? cf[3] : cf[0];
            st[ID] = q.d;
            st[IU] = q.u * q.d;
            st[IV] = q.v * q.d;
            st[IP] = q.p / (p.gamma0 - 1.0) + q.d * (q.u * q.u + q.v * q.v) * 0.5;
          } else {  // InitIsentropicVortexFunctor2D
            const double xp = x - pr.iv_x, yp = y - pr.iv_y, r2 = xp * xp + yp * yp;
            const double r = std::sqrt(r2);
            const double du = -yp * pr.iv_beta / (2 * PI) * std::exp(0.5 * (1.0 - r * r));
            const double dv = xp * pr.iv_beta / (2 * PI) * std::exp(0.5 * (1.0 - r * r));
            const double T = pr.iv_t_a - (p.gamma0 - 1) * pr.iv_beta * pr.iv_beta /
                                             (8 * p.gamma0 * PI * PI) * std::exp(1.0 - r * r);
            const double rho = pr.iv_rho_a * std::pow(T / pr.iv_t_a, 1.0 / (p.gamma0 - 1));
            st[ID] = rho;
            st[IU] = rho * (pr.iv_u_a + du);
            st[IV] = rho * (pr.iv_v_a + dv);
            st[IP] = rho * T / (p.gamma0 - 1.0) + 0.5 * rho * (pr.iv_u_a + du) * (pr.iv_u_a + du) +
                     0.5 * rho * (pr.iv_v_a + dv) * (pr.iv_v_a + dv);
          }
          for (int n = 0; n < g.nv; ++n) U[g.at(i, j, k, n)] = st[n];
          continue;
        }
        if (pr.kind == PPK_PROBLEM_BLAST) {
          double d2 = (x - pr.blast_center_x) * (x - pr.blast_center_x) +
                      (y - pr.blast_center_y) * (y - pr.blast_center_y);
          if (g.dim == 3) d2 += (z - pr.blast_center_z) * (z - pr.blast_center_z);
          const bool in = d2 < pr.blast_radius * pr.blast_radius;
          d = in ? pr.blast_density_in : pr.blast_density_out;
          e = (in ? pr.blast_pressure_in : pr.blast_pressure_out) / (p.gamma0 - 1.0);
          u = v = w = 0.0;
        } else {  // implode
          bool out;
          if (g.dim == 2)
            out = pr.implode_shape == 1 ? (x + y * y > 0.5 && x + y * y < 1.5)
                                        : (x + y > (p.xmin + p.xmax) / 2. + p.ymin);
          else
            out = pr.implode_shape == 1 ? (x + y + z > 0.5 && x + y + z < 2.5)
                                        : (x + y + z > (p.xmin + p.xmax) / 2. + p.ymin + p.zmin);
          const double *s = out ? pr.implode_out : pr.implode_in;
          d = s[0];
          u = s[2];
          v = s[3];
          w = s[4];
          e = g.dim == 2 ? s[1] / (p.gamma0 - 1.0) + 0.5 * d * (u * u + v * v)
                         : s[1] / (p.gamma0 - 1.0) + 0.5 * d * (u * u + v * v + w * w);
        }
        U[g.at(i, j, k, ID)] = d;
        U[g.at(i, j, k, IP)] = e;
        U[g.at(i, j, k, IU)] = u;  // the reference stores the velocity, not the momentum
        U[g.at(i, j, k, IV)] = v;
        if (g.dim == 3) U[g.at(i, j, k, IW)] = w;
      }
}

}  // namespace

extern "C" {

// OpenMP threads of the parallel loops (the CPU baseline records it)
int hora_threads() {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

struct hora_solver {
  HydroSolver s;
  explicit hora_solver(const ppk_mhd_params &p) : s(p) {}
};

hora_solver *hora_create(const ppk_mhd_params *p) { return new hora_solver(*p); }
void hora_destroy(hora_solver *s) { delete s; }
void hora_init_problem(hora_solver *s, const ppk_mhd_problem *pr) {
  init(s->s.p, *pr, s->s.U.data());
  gravity_field(s->s.p, *pr, s->s.G.data());
  boundaries(s->s.p, s->s.U.data());
  s->s.U2 = s->s.U;
  s->s.t = 0.0;
  s->s.it = 0;
}
void hora_init_array(const ppk_mhd_params *p, const ppk_mhd_problem *pr, double *U) {
  init(*p, *pr, U);
}
void hora_gravity_field(const ppk_mhd_params *p, const ppk_mhd_problem *pr, double *G) {
  gravity_field(*p, *pr, G);
}
void hora_set_state(hora_solver *s, const double *U) {
  std::vector<double> &cur = (s->s.it % 2 == 0) ? s->s.U : s->s.U2;
  std::memcpy(cur.data(), U, sizeof(double) * cur.size());
}
void hora_get_state(hora_solver *s, double *U) {
  const std::vector<double> &cur = (s->s.it % 2 == 0) ? s->s.U : s->s.U2;
  std::memcpy(U, cur.data(), sizeof(double) * cur.size());
}
void hora_step(hora_solver *s, double dt, int mode) { s->s.step(dt, mode); }
void hora_get_time(hora_solver *s, double *t, double *dt, long long *it) {
  *t = s->s.t;
  *dt = s->s.dt;
  *it = s->s.it;
}
void hora_set_time(hora_solver *s, double t, long long it) {
  // keep the current data in the current buffer
  if ((it % 2) != (s->s.it % 2)) std::swap(s->s.U, s->s.U2);
  s->s.t = t;
  s->s.it = it;
}
void hora_boundaries(const ppk_mhd_params *p, double *U) { boundaries(*p, U); }
void hora_riemann(const ppk_mhd_params *p, const double *ql, const double *qr, double *f) {
  const int nv = p->dim == 3 ? 5 : 4;
  S5 a, b, o;
  for (int v = 0; v < nv; ++v) {
    a[v] = ql[v];
    b[v] = qr[v];
  }
  solve(*p, nv, a, b, o);
  for (int v = 0; v < nv; ++v) f[v] = D(o[v]);
}
double hora_inv_dt(hora_solver *s) {
  const std::vector<double> &cur = (s->s.it % 2 == 0) ? s->s.U : s->s.U2;
  return D(s->s.inv_dt(cur));
}

/* operation counts of the steps since the last reset (HORA_COUNT build; zeros otherwise),
   per stage: counts[5 * stage + {0 add/sub, 1 mul, 2 div, 3 sqrt, 4 min/max/abs}] */
int hora_op_stages(void) {
#ifdef HORA_COUNT
  return kStages;
#else
  return 0;
#endif
}
void hora_op_counts(long long *counts) {
#ifdef HORA_COUNT
  for (int st = 0; st < kStages; ++st) {
    const OpCounts &o = g_stage_ops[st];
    const long long v[5] = {o.add, o.mul, o.div, o.sqrt, o.sel};
    for (int m = 0; m < 5; ++m) counts[5 * st + m] = v[m];
  }
#else
  (void)counts;
#endif
}
void hora_op_counts_reset(void) {
#ifdef HORA_COUNT
  for (auto &o : g_stage_ops) o = OpCounts{};
#endif
}

}  // extern "C"
