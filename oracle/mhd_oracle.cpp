// TEST INFRASTRUCTURE -- NOT PART OF THE PRODUCT.
//
// CPU oracle: restatement of ppkMHD's MHD MUSCL-Hancock step, implementation version 0 (the
// only deterministic one: versions 1 and 2 scatter with atomics), 2D and 3D.  Used by tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg as the checker, never by the product.
//
// Every function cites the reference file:line it restates (paths relative to the reference
// repository root).  Arithmetic follows the reference expression by expression (same operand
// order, same grouping) so that, built with -ffp-contract=off, it is a bit-level statement of
// the reference algorithm on an IEEE-754 double machine.
//
// Built twice from this one source (oracle/Makefile):
//   liboracle.so        real = double, OpenMP loops (also the CPU baseline);
//   liboracle_count.so  real = an instrumented scalar that counts add/mul/div/sqrt/select,
//                       giving the per-cell operation count of the step.

#include "mhd_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#ifdef _OPENMP
#include <omp.h>
#endif
#include <vector>

#ifdef ORA_COUNT
#define ORA_COUNTING 1
#else
#define ORA_COUNTING 0
#endif

namespace {

// ---------------------------------------------------------------------------------------------
// scalar type
// ---------------------------------------------------------------------------------------------
#ifdef ORA_COUNT
struct OpCounts {
  long long add = 0, mul = 0, div = 0, sqrt = 0, sel = 0;
};
// counts per stage of the step (index set by the solver before each stage)
constexpr int kStages = 10;
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
inline real operator+(real a) { return a; }
inline real &operator+=(real &a, real b) { a = a + b; return a; }
inline real &operator-=(real &a, real b) { a = a - b; return a; }
inline real &operator*=(real &a, real b) { a = a * b; return a; }
inline bool operator<(real a, real b) { return a.v < b.v; }
inline bool operator>(real a, real b) { return a.v > b.v; }
inline bool operator<=(real a, real b) { return a.v <= b.v; }
inline bool operator>=(real a, real b) { return a.v >= b.v; }
inline real SQRT(real a) { ++g_ops.sqrt; return real(std::sqrt(a.v)); }
inline real FABS(real a) { ++g_ops.sel; return real(std::fabs(a.v)); }
inline real FMAX(real a, real b) { ++g_ops.sel; return real(std::fmax(a.v, b.v)); }
inline real FMIN(real a, real b) { ++g_ops.sel; return real(std::fmin(a.v, b.v)); }
inline real COPYSIGN(real a, real b) { return real(std::copysign(a.v, b.v)); }
inline double D(real a) { return a.v; }
#define ORA_STAGE(n) (g_stage = (n))
#else
#define ORA_STAGE(n) ((void)0)
typedef double real;
inline double SQRT(double a) { return std::sqrt(a); }
inline double FABS(double a) { return std::fabs(a); }
inline double FMAX(double a, double b) { return std::fmax(a, b); }
inline double FMIN(double a, double b) { return std::fmin(a, b); }
inline double COPYSIGN(double a, double b) { return std::copysign(a, b); }
inline double D(double a) { return a; }
#endif

// variable indexes, src/shared/enums.h:17-34
enum { ID = 0, IP = 1, IU = 2, IV = 3, IW = 4, IA = 5, IB = 6, IC = 7 };
enum { IX = 0, IY = 1, IZ = 2 };
// edge indexes, src/shared/enums.h:184-198
enum { IRT = 0, IRB = 1, ILT = 2, ILB = 3 };
enum { ILL = 0, IRL = 1, ILR = 2, IRR = 3 };
// EMF slots, src/shared/enums.h:208-217
enum { EMFX = 0, EMFY = 1, EMFZ = 2 };
enum { I_EMFZ = 0, I_EMFY = 1, I_EMFX = 2 };

struct St {
  real v[8];
  real &operator[](int i) { return v[i]; }
  const real &operator[](int i) const { return v[i]; }
};

// ---------------------------------------------------------------------------------------------
// padded box and array access
// ---------------------------------------------------------------------------------------------
struct Box {
  int nx, ny, nz, gw, is, js, ks, dim;
  explicit Box(const ppk_mhd_params &p) {
    dim = p.dim;
    gw = p.ghost_width;
    nx = p.nx;
    ny = p.ny;
    nz = dim == 3 ? p.nz : 1;
    // src/shared/HydroParams.cpp:426-436
    is = nx + 2 * gw;
    js = ny + 2 * gw;
    ks = dim == 3 ? nz + 2 * gw : 1;
  }
  long long cells() const { return (long long)is * js * ks; }
  long long at(int i, int j, int k, int v) const {
    return (((long long)v * ks + k) * js + j) * (long long)is + i;
  }
};

inline void get_state(const Box &b, const double *a, int i, int j, int k, St &q) {
  for (int v = 0; v < 8; ++v) q[v] = real(a[b.at(i, j, k, v)]);
}
inline void set_state(const Box &b, double *a, int i, int j, int k, const St &q) {
  for (int v = 0; v < 8; ++v) a[b.at(i, j, k, v)] = D(q[v]);
}
inline void swap_values(real &a, real &b) {
  real t = a;
  a = b;
  b = t;
}

// ---------------------------------------------------------------------------------------------
// MHD utilities, src/shared/mhd_utils.h
// ---------------------------------------------------------------------------------------------
// FMAX4 / FMIN4 / FMAX5 are compare-and-select, not fmax: src/shared/mhd_utils.h:36-77
inline real max4(real a0, real a1, real a2, real a3) {
  real r = a0;
  r = (a1 > r) ? a1 : r;
  r = (a2 > r) ? a2 : r;
  r = (a3 > r) ? a3 : r;
  return r;
}
inline real min4(real a0, real a1, real a2, real a3) {
  real r = a0;
  r = (a1 < r) ? a1 : r;
  r = (a2 < r) ? a2 : r;
  r = (a3 < r) ? a3 : r;
  return r;
}
inline real max5(real a0, real a1, real a2, real a3, real a4) {
  real r = a0;
  r = (a1 > r) ? a1 : r;
  r = (a2 > r) ? a2 : r;
  r = (a3 > r) ? a3 : r;
  r = (a4 > r) ? a4 : r;
  return r;
}

// find_speed_fast<dir>, src/shared/mhd_utils.h:89-117
inline real speed_fast(const St &q, int dir, real gamma0) {
  real d = q[ID], p = q[IP], a = q[IA], b = q[IB], c = q[IC];
  real b2 = a * a + b * b + c * c;
  real c2 = gamma0 * p / d;
  real d2 = 0.5 * (b2 / d + c2);
  real bn = dir == IX ? a : (dir == IY ? b : c);
  return SQRT(d2 + SQRT(d2 * d2 - c2 * bn * bn / d));
}

// find_speed_info<NDIM> summed into the dt reduction, src/shared/mhd_utils.h:319-366
inline real inv_dt_cell(const St &q, const ppk_mhd_params &p) {
  real gamma0 = p.gamma0;
  real u = q[IU], v = q[IV], w = q[IW];
  real d = q[ID], pp = q[IP], a = q[IA], b = q[IB], c = q[IC];
  real b2 = a * a + b * b + c * c;
  real c2 = gamma0 * pp / d;
  real d2 = 0.5 * (b2 / d + c2);
  real cf = SQRT(d2 + SQRT(d2 * d2 - c2 * a * a / d));
  real vx = cf + FABS(u);
  cf = SQRT(d2 + SQRT(d2 * d2 - c2 * b * b / d));
  real vy = cf + FABS(v);
  if (p.dim == 3) {
    cf = SQRT(d2 + SQRT(d2 * d2 - c2 * c * c / d));
    real vz = cf + FABS(w);
    // src/muscl/MHDRunFunctors3D.h:76
    return vx / real(p.dx) + vy / real(p.dy) + vz / real(p.dz);
  }
  // src/muscl/MHDRunFunctors2D.h:73
  return vx / real(p.dx) + vy / real(p.dy);
}

// find_mhd_flux, src/shared/mhd_utils.h:173-231 (used by HLL / LLF)
inline void mhd_flux(const St &q, St &cvar, St &ff, const ppk_mhd_params &par) {
  real gamma0 = par.gamma0;
  real p;
  if (par.cIso > 0) {
    p = q[ID] * real(par.cIso) * real(par.cIso);
  } else {
    p = q[IP];
  }
  real entho = 1.0 / (gamma0 - 1.0);
  real d = q[ID], u = q[IU], v = q[IV], w = q[IW], a = q[IA], b = q[IB], c = q[IC];
  real ecin = 0.5 * (u * u + v * v + w * w) * d;
  real emag = 0.5 * (a * a + b * b + c * c);
  real etot = p * entho + ecin + emag;
  real ptot = p + emag;
  cvar[ID] = d;
  cvar[IP] = etot;
  cvar[IU] = d * u;
  cvar[IV] = d * v;
  cvar[IW] = d * w;
  cvar[IA] = a;
  cvar[IB] = b;
  cvar[IC] = c;
  ff[ID] = d * u;
  ff[IP] = (etot + ptot) * u - a * (a * u + b * v + c * w);
  ff[IU] = d * u * u - a * a + ptot;
  ff[IV] = d * u * v - a * b;
  ff[IW] = d * u * w - a * c;
  ff[IA] = 0.0;
  ff[IB] = b * u - a * v;
  ff[IC] = c * u - a * w;
}

// ---------------------------------------------------------------------------------------------
// Riemann solvers, src/shared/RiemannSolvers_MHD.h
// ---------------------------------------------------------------------------------------------
// riemann_hll, src/shared/RiemannSolvers_MHD.h:27-67
void riemann_hll(St &ql, St &qr, St &flux, const ppk_mhd_params &p) {
  real bx_mean = 0.5 * (ql[IA] + qr[IA]);
  ql[IA] = bx_mean;
  qr[IA] = bx_mean;
  St ul, fl, ur, fr;
  mhd_flux(ql, ul, fl, p);
  mhd_flux(qr, ur, fr, p);
  real cfl = speed_fast(ql, IX, p.gamma0);
  real cfr = speed_fast(qr, IX, p.gamma0);
  real vl = ql[IU], vr = qr[IU];
  real sl = FMIN(FMIN(vl, vr) - FMAX(cfl, cfr), 0.0);
  real sr = FMAX(FMAX(vl, vr) + FMAX(cfl, cfr), 0.0);
  for (int v = 0; v < 8; ++v)
    flux[v] = (sr * fl[v] - sl * fr[v] + sr * sl * (ur[v] - ul[v])) / (sr - sl);
}

// riemann_llf, src/shared/RiemannSolvers_MHD.h:81-111
void riemann_llf(St &ql, St &qr, St &flux, const ppk_mhd_params &p) {
  real bx_mean = 0.5 * (ql[IA] + qr[IA]);
  ql[IA] = bx_mean;
  qr[IA] = bx_mean;
  St ul, fl, ur, fr;
  mhd_flux(ql, ul, fl, p);
  mhd_flux(qr, ur, fr, p);
  for (int v = 0; v < 8; ++v) flux[v] = (fl[v] + fr[v]) / 2.0;
  // find_speed_info (1D version), src/shared/mhd_utils.h:377-403
  auto info = [&](const St &q) {
    real d = q[ID], pp = q[IP], a = q[IA], b = q[IB], c = q[IC];
    real b2 = a * a + b * b + c * c;
    real c2 = real(p.gamma0) * pp / d;
    real d2 = 0.5 * (b2 / d + c2);
    real cf = SQRT(d2 + SQRT(d2 * d2 - c2 * a * a / d));
    return cf + FABS(q[IU]);
  };
  real vel_info = FMAX(info(ql), info(qr));
  for (int v = 0; v < 8; ++v) flux[v] = flux[v] - vel_info * (ur[v] - ul[v]) / 2.0;
}

// riemann_hlld (Miyoshi & Kusano 2005 as in RAMSES/DUMSES),
// src/shared/RiemannSolvers_MHD.h:133-367
void riemann_hlld(St &qleft, St &qright, St &flux, const ppk_mhd_params &par) {
  const real gamma0 = par.gamma0;
  const real entho = 1.0 / (gamma0 - 1.0);

  real a = 0.5 * (qleft[IA] + qright[IA]);
  real sgnm = (a >= 0) ? 1.0 : -1.0;
  qleft[IA] = a;
  qright[IA] = a;
  if (par.cIso > 0) {
    qleft[IP] = qleft[ID] * real(par.cIso) * real(par.cIso);
    qright[IP] = qright[ID] * real(par.cIso) * real(par.cIso);
  }

  real rl = qleft[ID], pl = qleft[IP], ul = qleft[IU], vl = qleft[IV], wl = qleft[IW];
  real bl = qleft[IB], cl = qleft[IC];
  real ecinl = 0.5 * (ul * ul + vl * vl + wl * wl) * rl;
  real emagl = 0.5 * (a * a + bl * bl + cl * cl);
  real etotl = pl * entho + ecinl + emagl;
  real ptotl = pl + emagl;
  real vdotbl = ul * a + vl * bl + wl * cl;

  real rr = qright[ID], pr = qright[IP], ur = qright[IU], vr = qright[IV], wr = qright[IW];
  real br = qright[IB], cr = qright[IC];
  real ecinr = 0.5 * (ur * ur + vr * vr + wr * wr) * rr;
  real emagr = 0.5 * (a * a + br * br + cr * cr);
  real etotr = pr * entho + ecinr + emagr;
  real ptotr = pr + emagr;
  real vdotbr = ur * a + vr * br + wr * cr;

  real cfastl = speed_fast(qleft, IX, gamma0);
  real cfastr = speed_fast(qright, IX, gamma0);

  real sl = FMIN(ul, ur) - FMAX(cfastl, cfastr);
  real sr = FMAX(ul, ur) + FMAX(cfastl, cfastr);

  real rcl = rl * (ul - sl);
  real rcr = rr * (sr - ur);

  real ustar = (rcr * ur + rcl * ul + (ptotl - ptotr)) / (rcr + rcl);
  real ptotstar = (rcr * ptotl + rcl * ptotr + rcl * rcr * (ul - ur)) / (rcr + rcl);

  // left star region
  real rstarl = rl * (sl - ul) / (sl - ustar);
  real estar = rl * (sl - ul) * (sl - ustar) - a * a;
  real el = rl * (sl - ul) * (sl - ul) - a * a;
  real vstarl, wstarl, bstarl, cstarl;
  if (a * a > 0 && FABS(estar / (a * a) - 1.0) <= 1e-8) {
    vstarl = vl;
    bstarl = bl;
    wstarl = wl;
    cstarl = cl;
  } else {
    vstarl = vl - a * bl * (ustar - ul) / estar;
    bstarl = bl * el / estar;
    wstarl = wl - a * cl * (ustar - ul) / estar;
    cstarl = cl * el / estar;
  }
  real vdotbstarl = ustar * a + vstarl * bstarl + wstarl * cstarl;
  real etotstarl =
      ((sl - ul) * etotl - ptotl * ul + ptotstar * ustar + a * (vdotbl - vdotbstarl)) /
      (sl - ustar);
  real sqrrstarl = SQRT(rstarl);
  real calfvenl = FABS(a) / sqrrstarl;
  real sal = ustar - calfvenl;

  // right star region
  real rstarr = rr * (sr - ur) / (sr - ustar);
  estar = rr * (sr - ur) * (sr - ustar) - a * a;
  real er = rr * (sr - ur) * (sr - ur) - a * a;
  real vstarr, wstarr, bstarr, cstarr;
  if (a * a > 0 && FABS(estar / (a * a) - 1.0) <= 1e-8) {
    vstarr = vr;
    bstarr = br;
    wstarr = wr;
    cstarr = cr;
  } else {
    vstarr = vr - a * br * (ustar - ur) / estar;
    bstarr = br * er / estar;
    wstarr = wr - a * cr * (ustar - ur) / estar;
    cstarr = cr * er / estar;
  }
  real vdotbstarr = ustar * a + vstarr * bstarr + wstarr * cstarr;
  real etotstarr =
      ((sr - ur) * etotr - ptotr * ur + ptotstar * ustar + a * (vdotbr - vdotbstarr)) /
      (sr - ustar);
  real sqrrstarr = SQRT(rstarr);
  real calfvenr = FABS(a) / sqrrstarr;
  real sar = ustar + calfvenr;

  // double star region
  real vstarstar =
      (sqrrstarl * vstarl + sqrrstarr * vstarr + sgnm * (bstarr - bstarl)) / (sqrrstarl + sqrrstarr);
  real wstarstar =
      (sqrrstarl * wstarl + sqrrstarr * wstarr + sgnm * (cstarr - cstarl)) / (sqrrstarl + sqrrstarr);
  real bstarstar = (sqrrstarl * bstarr + sqrrstarr * bstarl +
                    sgnm * sqrrstarl * sqrrstarr * (vstarr - vstarl)) /
                   (sqrrstarl + sqrrstarr);
  real cstarstar = (sqrrstarl * cstarr + sqrrstarr * cstarl +
                    sgnm * sqrrstarl * sqrrstarr * (wstarr - wstarl)) /
                   (sqrrstarl + sqrrstarr);
  real vdotbstarstar = ustar * a + vstarstar * bstarstar + wstarstar * cstarstar;
  real etotstarstarl = etotstarl - sgnm * sqrrstarl * (vdotbstarl - vdotbstarstar);
  real etotstarstarr = etotstarr + sgnm * sqrrstarr * (vdotbstarr - vdotbstarstar);

  // sample at x/t = 0
  real ro, uo, vo, wo, bo, co, ptoto, etoto, vdotbo;
  if (sl > 0) {
    ro = rl; uo = ul; vo = vl; wo = wl; bo = bl; co = cl;
    ptoto = ptotl; etoto = etotl; vdotbo = vdotbl;
  } else if (sal > 0) {
    ro = rstarl; uo = ustar; vo = vstarl; wo = wstarl; bo = bstarl; co = cstarl;
    ptoto = ptotstar; etoto = etotstarl; vdotbo = vdotbstarl;
  } else if (ustar > 0) {
    ro = rstarl; uo = ustar; vo = vstarstar; wo = wstarstar; bo = bstarstar; co = cstarstar;
    ptoto = ptotstar; etoto = etotstarstarl; vdotbo = vdotbstarstar;
  } else if (sar > 0) {
    ro = rstarr; uo = ustar; vo = vstarstar; wo = wstarstar; bo = bstarstar; co = cstarstar;
    ptoto = ptotstar; etoto = etotstarstarr; vdotbo = vdotbstarstar;
  } else if (sr > 0) {
    ro = rstarr; uo = ustar; vo = vstarr; wo = wstarr; bo = bstarr; co = cstarr;
    ptoto = ptotstar; etoto = etotstarr; vdotbo = vdotbstarr;
  } else {
    ro = rr; uo = ur; vo = vr; wo = wr; bo = br; co = cr;
    ptoto = ptotr; etoto = etotr; vdotbo = vdotbr;
  }

  flux[ID] = ro * uo;
  flux[IP] = (etoto + ptoto) * uo - a * vdotbo;
  flux[IU] = ro * uo * uo - a * a + ptoto;
  flux[IV] = ro * uo * vo - a * bo;
  flux[IW] = ro * uo * wo - a * co;
  flux[IA] = 0.0;
  flux[IB] = bo * uo - a * vo;
  flux[IC] = co * uo - a * wo;
}

// riemann_mhd dispatch, src/shared/RiemannSolvers_MHD.h:372-392
void riemann_mhd(St &ql, St &qr, St &flux, const ppk_mhd_params &p) {
  if (p.riemann == PPK_RIEMANN_HLLD)
    riemann_hlld(ql, qr, flux, p);
  else if (p.riemann == PPK_RIEMANN_HLL)
    riemann_hll(ql, qr, flux, p);
  else if (p.riemann == PPK_RIEMANN_LLF)
    riemann_llf(ql, qr, flux, p);
  else
    for (int v = 0; v < 8; ++v) flux[v] = 0.0;
}

// mag_riemann2d_hlld, src/shared/RiemannSolvers_MHD.h:398-630.  The emf is blended from the
// five regions with integer 0/1 masks built from copysign (so a signed zero selects), as the
// reference does (:569-596).
real mag_riemann2d_hlld(const St (&q)[4], const real (&e)[4], const ppk_mhd_params &p) {
  const St &qLL = q[ILL];
  const St &qRL = q[IRL];
  const St &qLR = q[ILR];
  const St &qRR = q[IRR];
  const real ELL = e[ILL], ERL = e[IRL], ELR = e[ILR], ERR = e[IRR];
  const real gamma0 = p.gamma0;

  real rLL = qLL[ID], uLL = qLL[IU], vLL = qLL[IV], aLL = qLL[IA], bLL = qLL[IB], cLL = qLL[IC];
  real pLL = qLL[IP];
  real rLR = qLR[ID], uLR = qLR[IU], vLR = qLR[IV], aLR = qLR[IA], bLR = qLR[IB], cLR = qLR[IC];
  real pLR = qLR[IP];
  real rRL = qRL[ID], uRL = qRL[IU], vRL = qRL[IV], aRL = qRL[IA], bRL = qRL[IB], cRL = qRL[IC];
  real pRL = qRL[IP];
  real rRR = qRR[ID], uRR = qRR[IU], vRR = qRR[IV], aRR = qRR[IA], bRR = qRR[IB], cRR = qRR[IC];
  real pRR = qRR[IP];

  real cFastLLx = speed_fast(qLL, IX, gamma0);
  real cFastLRx = speed_fast(qLR, IX, gamma0);
  real cFastRLx = speed_fast(qRL, IX, gamma0);
  real cFastRRx = speed_fast(qRR, IX, gamma0);
  real cFastLLy = speed_fast(qLL, IY, gamma0);
  real cFastLRy = speed_fast(qLR, IY, gamma0);
  real cFastRLy = speed_fast(qRL, IY, gamma0);
  real cFastRRy = speed_fast(qRR, IY, gamma0);

  real SL = min4(uLL, uLR, uRL, uRR) - max4(cFastLLx, cFastLRx, cFastRLx, cFastRRx);
  real SR = max4(uLL, uLR, uRL, uRR) + max4(cFastLLx, cFastLRx, cFastRLx, cFastRRx);
  real SB = min4(vLL, vLR, vRL, vRR) - max4(cFastLLy, cFastLRy, cFastRLy, cFastRRy);
  real ST = max4(vLL, vLR, vRL, vRR) + max4(cFastLLy, cFastLRy, cFastRLy, cFastRRy);

  real PtotLL = pLL + 0.5 * (aLL * aLL + bLL * bLL + cLL * cLL);
  real PtotLR = pLR + 0.5 * (aLR * aLR + bLR * bLR + cLR * cLR);
  real PtotRL = pRL + 0.5 * (aRL * aRL + bRL * bRL + cRL * cRL);
  real PtotRR = pRR + 0.5 * (aRR * aRR + bRR * bRR + cRR * cRR);

  real rcLLx = rLL * (uLL - SL);
  real rcRLx = rRL * (SR - uRL);
  real rcLRx = rLR * (uLR - SL);
  real rcRRx = rRR * (SR - uRR);
  real rcLLy = rLL * (vLL - SB);
  real rcLRy = rLR * (ST - vLR);
  real rcRLy = rRL * (vRL - SB);
  real rcRRy = rRR * (ST - vRR);

  real ustar = (rcLLx * uLL + rcLRx * uLR + rcRLx * uRL + rcRRx * uRR +
                (PtotLL - PtotRL + PtotLR - PtotRR)) /
               (rcLLx + rcLRx + rcRLx + rcRRx);
  real vstar = (rcLLy * vLL + rcLRy * vLR + rcRLy * vRL + rcRRy * vRR +
                (PtotLL - PtotLR + PtotRL - PtotRR)) /
               (rcLLy + rcLRy + rcRLy + rcRRy);

  real rstarLLx = rLL * (SL - uLL) / (SL - ustar);
  real BstarLL = bLL * (SL - uLL) / (SL - ustar);
  real rstarLLy = rLL * (SB - vLL) / (SB - vstar);
  real AstarLL = aLL * (SB - vLL) / (SB - vstar);
  real rstarLL = rLL * (SL - uLL) / (SL - ustar) * (SB - vLL) / (SB - vstar);
  real EstarLLx = ustar * BstarLL - vLL * aLL;
  real EstarLLy = uLL * bLL - vstar * AstarLL;
  real EstarLL = ustar * BstarLL - vstar * AstarLL;

  real rstarLRx = rLR * (SL - uLR) / (SL - ustar);
  real BstarLR = bLR * (SL - uLR) / (SL - ustar);
  real rstarLRy = rLR * (ST - vLR) / (ST - vstar);
  real AstarLR = aLR * (ST - vLR) / (ST - vstar);
  real rstarLR = rLR * (SL - uLR) / (SL - ustar) * (ST - vLR) / (ST - vstar);
  real EstarLRx = ustar * BstarLR - vLR * aLR;
  real EstarLRy = uLR * bLR - vstar * AstarLR;
  real EstarLR = ustar * BstarLR - vstar * AstarLR;

  real rstarRLx = rRL * (SR - uRL) / (SR - ustar);
  real BstarRL = bRL * (SR - uRL) / (SR - ustar);
  real rstarRLy = rRL * (SB - vRL) / (SB - vstar);
  real AstarRL = aRL * (SB - vRL) / (SB - vstar);
  real rstarRL = rRL * (SR - uRL) / (SR - ustar) * (SB - vRL) / (SB - vstar);
  real EstarRLx = ustar * BstarRL - vRL * aRL;
  real EstarRLy = uRL * bRL - vstar * AstarRL;
  real EstarRL = ustar * BstarRL - vstar * AstarRL;

  real rstarRRx = rRR * (SR - uRR) / (SR - ustar);
  real BstarRR = bRR * (SR - uRR) / (SR - ustar);
  real rstarRRy = rRR * (ST - vRR) / (ST - vstar);
  real AstarRR = aRR * (ST - vRR) / (ST - vstar);
  real rstarRR = rRR * (SR - uRR) / (SR - ustar) * (ST - vRR) / (ST - vstar);
  real EstarRRx = ustar * BstarRR - vRR * aRR;
  real EstarRRy = uRR * bRR - vstar * AstarRR;
  real EstarRR = ustar * BstarRR - vstar * AstarRR;

  const real smallc = p.smallc;
  real calfvenL = max5(FABS(aLR) / SQRT(rstarLRx), FABS(AstarLR) / SQRT(rstarLR),
                       FABS(aLL) / SQRT(rstarLLx), FABS(AstarLL) / SQRT(rstarLL), smallc);
  real calfvenR = max5(FABS(aRR) / SQRT(rstarRRx), FABS(AstarRR) / SQRT(rstarRR),
                       FABS(aRL) / SQRT(rstarRLx), FABS(AstarRL) / SQRT(rstarRL), smallc);
  real calfvenB = max5(FABS(bLL) / SQRT(rstarLLy), FABS(BstarLL) / SQRT(rstarLL),
                       FABS(bRL) / SQRT(rstarRLy), FABS(BstarRL) / SQRT(rstarRL), smallc);
  real calfvenT = max5(FABS(bLR) / SQRT(rstarLRy), FABS(BstarLR) / SQRT(rstarLR),
                       FABS(bRR) / SQRT(rstarRRy), FABS(BstarRR) / SQRT(rstarRR), smallc);

  real SAL = FMIN(ustar - calfvenL, 0.0);
  real SAR = FMAX(ustar + calfvenR, 0.0);
  real SAB = FMIN(vstar - calfvenB, 0.0);
  real SAT = FMAX(vstar + calfvenT, 0.0);

  real AstarT = (SAR * AstarRR - SAL * AstarLR) / (SAR - SAL);
  real AstarB = (SAR * AstarRL - SAL * AstarLL) / (SAR - SAL);
  real BstarR = (SAT * BstarRR - SAB * BstarRL) / (SAT - SAB);
  real BstarL = (SAT * BstarLR - SAB * BstarLL) / (SAT - SAB);

  real E = 0.0, tmpE = 0.0;
  int SB_pos = (int)D(1.0 + COPYSIGN(1.0, SB)) / 2, SB_neg = 1 - SB_pos;
  int ST_pos = (int)D(1.0 + COPYSIGN(1.0, ST)) / 2, ST_neg = 1 - ST_pos;
  int SL_pos = (int)D(1.0 + COPYSIGN(1.0, SL)) / 2, SL_neg = 1 - SL_pos;
  int SR_pos = (int)D(1.0 + COPYSIGN(1.0, SR)) / 2, SR_neg = 1 - SR_pos;

  tmpE = (SAL * SAB * EstarRR - SAL * SAT * EstarRL - SAR * SAB * EstarLR + SAR * SAT * EstarLL) /
             (SAR - SAL) / (SAT - SAB) -
         SAT * SAB / (SAT - SAB) * (AstarT - AstarB) + SAR * SAL / (SAR - SAL) * (BstarR - BstarL);
  E += real(double(SB_neg * ST_pos * SL_neg * SR_pos)) * tmpE;

  tmpE = (SAR * EstarLLx - SAL * EstarRLx + SAR * SAL * (bRL - bLL)) / (SAR - SAL);
  tmpE = real(double(SL_pos)) * ELL + real(double(SL_neg * SR_neg)) * ERL +
         real(double(SL_neg * SR_pos)) * tmpE;
  E += real(double(SB_pos)) * tmpE;

  tmpE = (SAR * EstarLRx - SAL * EstarRRx + SAR * SAL * (bRR - bLR)) / (SAR - SAL);
  tmpE = real(double(SL_pos)) * ELR + real(double(SL_neg * SR_neg)) * ERR +
         real(double(SL_neg * SR_pos)) * tmpE;
  E += real(double(SB_neg * ST_neg)) * tmpE;

  tmpE = (SAT * EstarLLy - SAB * EstarLRy - SAT * SAB * (aLR - aLL)) / (SAT - SAB);
  E += real(double(SB_neg * ST_pos * SL_pos)) * tmpE;

  tmpE = (SAT * EstarRLy - SAB * EstarRRy - SAT * SAB * (aRR - aRL)) / (SAT - SAB);
  E += real(double(SB_neg * ST_pos * SL_neg * SR_neg)) * tmpE;

  return E;
}

// compute_emf<emfDir>, src/shared/RiemannSolvers_MHD.h:651-874.  qEdge holds RT, RB, LT, LB.
real compute_emf(int emfDir, const St (&qEdge)[4], const ppk_mhd_params &p) {
  const St &qRT = qEdge[IRT];
  const St &qLT = qEdge[ILT];
  const St &qRB = qEdge[IRB];
  const St &qLB = qEdge[ILB];
  St q[4];
  St &qLL = q[ILL];
  St &qRL = q[IRL];
  St &qLR = q[ILR];
  St &qRR = q[IRR];

  qLL[ID] = qRT[ID];
  qRL[ID] = qLT[ID];
  qLR[ID] = qRB[ID];
  qRR[ID] = qLB[ID];
  if (p.cIso > 0) {
    real c2 = real(p.cIso) * real(p.cIso);
    qLL[IP] = qLL[ID] * real(p.cIso) * real(p.cIso);
    qRL[IP] = qRL[ID] * real(p.cIso) * real(p.cIso);
    qLR[IP] = qLR[ID] * real(p.cIso) * real(p.cIso);
    qRR[IP] = qRR[ID] * real(p.cIso) * real(p.cIso);
    (void)c2;
  } else {
    qLL[IP] = qRT[IP];
    qRL[IP] = qLT[IP];
    qLR[IP] = qRB[IP];
    qRR[IP] = qLB[IP];
  }
  // (first parallel velocity, second parallel velocity, orthogonal velocity) and the same
  // for B, per EMF direction: Z -> (u, v, w), Y -> (w, u, v), X -> (v, w, u)
  int iu, iv, iw, ia, ib, ic;
  if (emfDir == EMFZ) {
    iu = IU; iv = IV; iw = IW; ia = IA; ib = IB; ic = IC;
  } else if (emfDir == EMFY) {
    iu = IW; iv = IU; iw = IV; ia = IC; ib = IA; ic = IB;
  } else {
    iu = IV; iv = IW; iw = IU; ia = IB; ib = IC; ic = IA;
  }
  qLL[IU] = qRT[iu];
  qRL[IU] = qLT[iu];
  qLR[IU] = qRB[iu];
  qRR[IU] = qLB[iu];

  qLL[IV] = qRT[iv];
  qRL[IV] = qLT[iv];
  qLR[IV] = qRB[iv];
  qRR[IV] = qLB[iv];

  qLL[IA] = 0.5 * (qRT[ia] + qLT[ia]);
  qRL[IA] = 0.5 * (qRT[ia] + qLT[ia]);
  qLR[IA] = 0.5 * (qRB[ia] + qLB[ia]);
  qRR[IA] = 0.5 * (qRB[ia] + qLB[ia]);

  qLL[IB] = 0.5 * (qRT[ib] + qRB[ib]);
  qRL[IB] = 0.5 * (qLT[ib] + qLB[ib]);
  qLR[IB] = 0.5 * (qRT[ib] + qRB[ib]);
  qRR[IB] = 0.5 * (qLT[ib] + qLB[ib]);

  qLL[IW] = qRT[iw];
  qRL[IW] = qLT[iw];
  qLR[IW] = qRB[iw];
  qRR[IW] = qLB[iw];

  qLL[IC] = qRT[ic];
  qRL[IC] = qLT[ic];
  qLR[IC] = qRB[ic];
  qRR[IC] = qLB[ic];

  real e[4];
  e[ILL] = qLL[IU] * qLL[IB] - qLL[IV] * qLL[IA];
  e[IRL] = qRL[IU] * qRL[IB] - qRL[IV] * qRL[IA];
  e[ILR] = qLR[IU] * qLR[IB] - qLR[IV] * qLR[IA];
  e[IRR] = qRR[IU] * qRR[IB] - qRR[IV] * qRR[IA];
  return mag_riemann2d_hlld(q, e, p);
}

// ---------------------------------------------------------------------------------------------
// cell-wise helpers, src/muscl/MHDBaseFunctor{2D,3D}.h
// ---------------------------------------------------------------------------------------------
// constoprim_mhd, src/muscl/MHDBaseFunctor2D.h:192-242 (= MHDBaseFunctor3D.h:192-242)
inline void constoprim(const St &u, const real (&mag)[3], St &q, const ppk_mhd_params &p) {
  real smallr = p.smallr;
  q[ID] = FMAX(u[ID], smallr);
  q[IU] = u[IU] / q[ID];
  q[IV] = u[IV] / q[ID];
  q[IW] = u[IW] / q[ID];
  q[IA] = 0.5 * (u[IA] + mag[0]);
  q[IB] = 0.5 * (u[IB] + mag[1]);
  q[IC] = 0.5 * (u[IC] + mag[2]);
  real eken = 0.5 * (q[IU] * q[IU] + q[IV] * q[IV] + q[IW] * q[IW]);
  real emag = 0.5 * (q[IA] * q[IA] + q[IB] * q[IB] + q[IC] * q[IC]);
  if (p.cIso > 0) {
    q[IP] = q[ID] * real(p.cIso) * real(p.cIso);
  } else {
    real eint = (u[IP] - emag) / q[ID] - eken;
    q[IP] = FMAX((real(p.gamma0) - 1.0) * q[ID] * eint, q[ID] * real(p.smallp));
  }
}

// limited slope (minmod for slope_type 1, MC for 2),
// src/muscl/MHDBaseFunctor2D.h:310-327 (= :261-295, :484-507, MHDBaseFunctor3D.h:264-344)
inline real limited_slope(real st, real q, real qplus, real qminus) {
  real dlft = st * (q - qminus);
  real drgt = st * (qplus - q);
  real dcen = 0.5 * (qplus - qminus);
  real dsgn = (dcen >= 0.0) ? 1.0 : -1.0;
  real slop = FMIN(FABS(dlft), FABS(drgt));
  real dlim = slop;
  if ((dlft * drgt) <= 0.0) dlim = 0.0;
  return dsgn * FMIN(dlim, FABS(dcen));
}

// ---------------------------------------------------------------------------------------------
// boundaries, src/shared/BoundariesFunctors.h:534-1057 driven in the order X then Y then Z by
// src/shared/SolverBase.cpp:513-537; each face runs over the full transverse range
// ---------------------------------------------------------------------------------------------
void fill_face(const ppk_mhd_params &p, double *U, int face) {
  Box b(p);
  const int gw = b.gw;
  const int dir = face / 2;     // 0 x, 1 y, 2 z
  const bool upper = face % 2;  // FACE_*MAX
  const int n = dir == 0 ? b.nx : (dir == 1 ? b.ny : b.nz);
  const int bc = p.bc[face];
  const int vflip = IU + dir, bflip = IA + dir;  // Dirichlet sign flips (:592-599, :898-904)
  const int s1 = dir == 0 ? b.js : b.is;          // transverse extents
  const int s2 = dir == 2 ? b.js : b.ks;
#pragma omp parallel for collapse(2) if (!ORA_COUNTING)
  for (int t2 = 0; t2 < s2; ++t2)
    for (int t1 = 0; t1 < s1; ++t1)
      for (int g = 0; g < gw; ++g) {
        int c = upper ? g + n + gw : g;  // ghost index along dir
        int c0;
        if (bc == PPK_BC_DIRICHLET)
          c0 = upper ? 2 * n + 2 * gw - 1 - c : 2 * gw - 1 - c;
        else if (bc == PPK_BC_NEUMANN)
          c0 = upper ? n + gw - 1 : gw;
        else
          c0 = upper ? c - n : n + c;
        for (int v = 0; v < 8; ++v) {
          double sign = 1.0;
          if (bc == PPK_BC_DIRICHLET && (v == vflip || v == bflip)) sign = -1.0;
          int i, j, k, i0, j0, k0;
          if (dir == 0) { i = c; j = t1; k = t2; i0 = c0; j0 = j; k0 = k; }
          else if (dir == 1) { i = t1; j = c; k = t2; i0 = i; j0 = c0; k0 = k; }
          else { i = t1; j = t2; k = c; i0 = i; j0 = j; k0 = c0; }
          U[b.at(i, j, k, v)] = U[b.at(i0, j0, k0, v)] * sign;
        }
      }
}

void fill_boundaries(const ppk_mhd_params &p, double *U) {
  int nfaces = p.dim == 3 ? 6 : 4;
  for (int f = 0; f < nfaces; ++f) fill_face(p, U, f);
}

// ConvertToPrimitivesFunctor{2D,3D}_MHD, src/muscl/MHDRunFunctors2D.h:105-154,
// src/muscl/MHDRunFunctors3D.h:107-158: every cell but the last one in each direction
void primitives(const ppk_mhd_params &p, const double *U, double *Q) {
  Box b(p);
  const int kend = p.dim == 3 ? b.ks - 1 : 1;
#pragma omp parallel for collapse(2) if (!ORA_COUNTING)
  for (int k = 0; k < kend; ++k)
    for (int j = 0; j < b.js - 1; ++j)
      for (int i = 0; i < b.is - 1; ++i) {
        St u, q;
        get_state(b, U, i, j, k, u);
        real mag[3];
        mag[0] = U[b.at(i + 1, j, k, IA)];
        mag[1] = U[b.at(i, j + 1, k, IB)];
        mag[2] = p.dim == 3 ? real(U[b.at(i, j, k + 1, IC)]) : real(0.0);
        constoprim(u, mag, q, p);
        set_state(b, Q, i, j, k, q);
      }
}

// ComputeDtFunctor{2D,3D}_MHD: max over interior cells, src/muscl/MHDRunFunctors2D.h:38-76,
// src/muscl/MHDRunFunctors3D.h:37-79
double inv_dt(const ppk_mhd_params &p, const double *Q) {
  Box b(p);
  const int gw = b.gw;
  const int k0 = p.dim == 3 ? gw : 0, k1 = p.dim == 3 ? b.ks - gw : 1;
  double best = 0.0;
#pragma omp parallel for collapse(2) reduction(max : best) if (!ORA_COUNTING)
  for (int k = k0; k < k1; ++k)
    for (int j = gw; j < b.js - gw; ++j)
      for (int i = gw; i < b.is - gw; ++i) {
        St q;
        get_state(b, Q, i, j, k, q);
        double v = D(inv_dt_cell(q, p));
        best = std::fmax(best, v);
      }
  return best;
}

// ---------------------------------------------------------------------------------------------
// 2D trace, src/muscl/MHDRunFunctors2D.h:848-904 + trace_unsplit_mhd_2d
// src/muscl/MHDBaseFunctor2D.h:774-1109
// ---------------------------------------------------------------------------------------------
struct Trace2D {
  St qm[2], qp[2], qEdge[4];  // qEdge order RT, RB, LT, LB
};

void trace_cell_2d(const ppk_mhd_params &p, const double *U, const double *Q, int i, int j,
                   real dtdx, real dtdy, Trace2D &out) {
  Box b(p);
  St qNb[3][3];
  real bf[4][4][3];
  for (int di = 0; di < 3; ++di)
    for (int dj = 0; dj < 3; ++dj) get_state(b, Q, i + di - 1, j + dj - 1, 0, qNb[di][dj]);
  for (int di = 0; di < 4; ++di)
    for (int dj = 0; dj < 4; ++dj) {
      bf[di][dj][0] = U[b.at(i + di - 1, j + dj - 1, 0, IA)];
      bf[di][dj][1] = U[b.at(i + di - 1, j + dj - 1, 0, IB)];
      bf[di][dj][2] = U[b.at(i + di - 1, j + dj - 1, 0, IC)];
    }
  const int C = 1;
  const real smallR = p.smallr, smallp = p.smallp, gamma = p.gamma0;
  const real st = p.slope_type;
  const St &q = qNb[C][C];

  // corner electric field Ez[di][dj] (:816-838)
  real Ez[2][2];
  for (int di = 0; di < 2; ++di)
    for (int dj = 0; dj < 2; ++dj) {
      int cx = C + di, cy = C + dj;
      real u = 0.25 * (qNb[cx - 1][cy - 1][IU] + qNb[cx - 1][cy][IU] + qNb[cx][cy - 1][IU] +
                       qNb[cx][cy][IU]);
      real v = 0.25 * (qNb[cx - 1][cy - 1][IV] + qNb[cx - 1][cy][IV] + qNb[cx][cy - 1][IV] +
                       qNb[cx][cy][IV]);
      real A = 0.5 * (bf[cx][cy - 1][0] + bf[cx][cy][0]);
      real B = 0.5 * (bf[cx - 1][cy][1] + bf[cx][cy][1]);
      Ez[di][dj] = u * B - v * A;
    }
  const real ELL = Ez[0][0], ELR = Ez[0][1], ERL = Ez[1][0], ERR = Ez[1][1];

  real r = q[ID], pp = q[IP], u = q[IU], v = q[IV], w = q[IW], A = q[IA], B = q[IB], Cc = q[IC];
  real AL = bf[C][C][0], AR = bf[C + 1][C][0], BL = bf[C][C][1], BR = bf[C][C + 1][1];

  // slope_unsplit_hydro_2d (:345-390), slope_type 1 or 2 only
  St dqx, dqy;
  for (int v8 = 0; v8 < 8; ++v8) {
    dqx[v8] = limited_slope(st, q[v8], qNb[C + 1][C][v8], qNb[C - 1][C][v8]);
    dqy[v8] = limited_slope(st, q[v8], qNb[C][C + 1][v8], qNb[C][C - 1][v8]);
  }
  real drx = dqx[ID]; drx *= 0.5;
  real dpx = dqx[IP]; dpx *= 0.5;
  real dux = dqx[IU]; dux *= 0.5;
  real dvx = dqx[IV]; dvx *= 0.5;
  real dwx = dqx[IW]; dwx *= 0.5;
  real dCx = dqx[IC]; dCx *= 0.5;
  real dBx = dqx[IB]; dBx *= 0.5;
  real dry = dqy[ID]; dry *= 0.5;
  real dpy = dqy[IP]; dpy *= 0.5;
  real duy = dqy[IU]; duy *= 0.5;
  real dvy = dqy[IV]; dvy *= 0.5;
  real dwy = dqy[IW]; dwy *= 0.5;
  real dCy = dqy[IC]; dCy *= 0.5;
  real dAy = dqy[IA]; dAy *= 0.5;

  // slope_unsplit_mhd_2d (:454-509) at (i,j), (i+1,j), (i,j+1): transverse face-B slopes
  real dALy = 0.5 * limited_slope(st, bf[C][C][0], bf[C][C + 1][0], bf[C][C - 1][0]);
  real dBLx = 0.5 * limited_slope(st, bf[C][C][1], bf[C + 1][C][1], bf[C - 1][C][1]);
  real dARy = 0.5 * limited_slope(st, bf[C + 1][C][0], bf[C + 1][C + 1][0], bf[C + 1][C - 1][0]);
  real dBRx = 0.5 * limited_slope(st, bf[C][C + 1][1], bf[C + 1][C + 1][1], bf[C - 1][C + 1][1]);

  real dAx = 0.5 * (AR - AL);
  real dBy = 0.5 * (BR - BL);

  // source terms (:969-995)
  real sr0 = (-u * drx - dux * r) * dtdx + (-v * dry - dvy * r) * dtdy;
  real su0 = (-u * dux - dpx / r - B * dBx / r - Cc * dCx / r) * dtdx + (-v * duy + B * dAy / r) * dtdy;
  real sv0 = (-u * dvx + A * dBx / r) * dtdx + (-v * dvy - dpy / r - A * dAy / r - Cc * dCy / r) * dtdy;
  real sw0 = (-u * dwx + A * dCx / r) * dtdx + (-v * dwy + B * dCy / r) * dtdy;
  real sp0 = (-u * dpx - dux * gamma * pp) * dtdx + (-v * dpy - dvy * gamma * pp) * dtdy;
  real sA0 = (u * dBy + B * duy - v * dAy - A * dvy) * dtdy;
  real sB0 = (-u * dBx - B * dux + v * dAx + A * dvx) * dtdx;
  real sC0 = (w * dAx + A * dwx - u * dCx - Cc * dux) * dtdx +
             (-v * dCy - Cc * dvy + w * dBy + B * dwy) * dtdy;
  real sAL0 = +(ELR - ELL) * 0.5 * dtdy;
  real sAR0 = +(ERR - ERL) * 0.5 * dtdy;
  real sBL0 = -(ERL - ELL) * 0.5 * dtdx;
  real sBR0 = -(ERR - ELR) * 0.5 * dtdx;

  r = r + sr0; u = u + su0; v = v + sv0; w = w + sw0; pp = pp + sp0;
  A = A + sA0; B = B + sB0; Cc = Cc + sC0;
  AL = AL + sAL0; AR = AR + sAR0; BL = BL + sBL0; BR = BR + sBR0;

  // face and edge states with density and pressure floors (pressure floor smallp*rho in 2D)
  auto floor2 = [&](St &s) {
    s[ID] = FMAX(smallR, s[ID]);
    s[IP] = FMAX(smallp * s[ID], s[IP]);
  };
  St *qp = out.qp, *qm = out.qm;
  qp[0][ID] = r - drx; qp[0][IU] = u - dux; qp[0][IV] = v - dvx; qp[0][IW] = w - dwx;
  qp[0][IP] = pp - dpx; qp[0][IA] = AL; qp[0][IB] = B - dBx; qp[0][IC] = Cc - dCx;
  floor2(qp[0]);
  qm[0][ID] = r + drx; qm[0][IU] = u + dux; qm[0][IV] = v + dvx; qm[0][IW] = w + dwx;
  qm[0][IP] = pp + dpx; qm[0][IA] = AR; qm[0][IB] = B + dBx; qm[0][IC] = Cc + dCx;
  floor2(qm[0]);
  qp[1][ID] = r - dry; qp[1][IU] = u - duy; qp[1][IV] = v - dvy; qp[1][IW] = w - dwy;
  qp[1][IP] = pp - dpy; qp[1][IA] = A - dAy; qp[1][IB] = BL; qp[1][IC] = Cc - dCy;
  floor2(qp[1]);
  qm[1][ID] = r + dry; qm[1][IU] = u + duy; qm[1][IV] = v + dvy; qm[1][IW] = w + dwy;
  qm[1][IP] = pp + dpy; qm[1][IA] = A + dAy; qm[1][IB] = BR; qm[1][IC] = Cc + dCy;
  floor2(qm[1]);

  St &qRT = out.qEdge[IRT], &qRB = out.qEdge[IRB], &qLT = out.qEdge[ILT], &qLB = out.qEdge[ILB];
  qRT[ID] = r + (+drx + dry); qRT[IU] = u + (+dux + duy); qRT[IV] = v + (+dvx + dvy);
  qRT[IW] = w + (+dwx + dwy); qRT[IP] = pp + (+dpx + dpy); qRT[IA] = AR + (+dARy);
  qRT[IB] = BR + (+dBRx); qRT[IC] = Cc + (+dCx + dCy);
  floor2(qRT);
  qRB[ID] = r + (+drx - dry); qRB[IU] = u + (+dux - duy); qRB[IV] = v + (+dvx - dvy);
  qRB[IW] = w + (+dwx - dwy); qRB[IP] = pp + (+dpx - dpy); qRB[IA] = AR + (-dARy);
  qRB[IB] = BL + (+dBLx); qRB[IC] = Cc + (+dCx - dCy);
  floor2(qRB);
  qLB[ID] = r + (-drx - dry); qLB[IU] = u + (-dux - duy); qLB[IV] = v + (-dvx - dvy);
  qLB[IW] = w + (-dwx - dwy); qLB[IP] = pp + (-dpx - dpy); qLB[IA] = AL + (-dALy);
  qLB[IB] = BL + (-dBLx); qLB[IC] = Cc + (-dCx - dCy);
  floor2(qLB);
  qLT[ID] = r + (-drx + dry); qLT[IU] = u + (-dux + duy); qLT[IV] = v + (-dvx + dvy);
  qLT[IW] = w + (-dwx + dwy); qLT[IP] = pp + (-dpx + dpy); qLT[IA] = AL + (+dALy);
  qLT[IB] = BR + (-dBRx); qLT[IC] = Cc + (-dCx + dCy);
  floor2(qLT);
}

// ---------------------------------------------------------------------------------------------
// 3D trace, src/muscl/MHDRunFunctors3D.h:669-843 + trace_unsplit_mhd_3d_simpler
// src/muscl/MHDBaseFunctor3D.h:688-1114; E field and magnetic slopes precomputed by
// ComputeElecFieldFunctor3D (:278-362) and ComputeMagSlopesFunctor3D (:441-538)
// ---------------------------------------------------------------------------------------------
struct Trace3D {
  St qm[3], qp[3];
  St qEdge[4][3];  // [RT,RB,LT,LB][X,Y,Z edge]
};

void trace_cell_3d(const ppk_mhd_params &p, const double *U, const double *Q, const double *E,
                   const double *dA, const double *dB, const double *dC, int i, int j, int k,
                   real dtdx, real dtdy, real dtdz, Trace3D &out) {
  Box b(p);
  const real st = p.slope_type;
  St q, qPX, qMX, qPY, qMY, qPZ, qMZ;
  get_state(b, Q, i, j, k, q);
  get_state(b, Q, i + 1, j, k, qPX);
  get_state(b, Q, i - 1, j, k, qMX);
  get_state(b, Q, i, j + 1, k, qPY);
  get_state(b, Q, i, j - 1, k, qMY);
  get_state(b, Q, i, j, k + 1, qPZ);
  get_state(b, Q, i, j, k - 1, qMZ);

  // slope_unsplit_hydro_3d (7-state form), src/muscl/MHDBaseFunctor3D.h:362-495.  The
  // reference passes qMinusY[IC] as the z-minus neighbour of IC (:458); dq[IZ][IC] is never
  // read by the trace (dCz comes from the faces), so the slip has no effect; kept as is.
  St dq[3];
  if (p.slope_type == 1.0 || p.slope_type == 2.0) {
    for (int v = 0; v < 8; ++v) {
      dq[IX][v] = limited_slope(st, q[v], qPX[v], qMX[v]);
      dq[IY][v] = limited_slope(st, q[v], qPY[v], qMY[v]);
      dq[IZ][v] = limited_slope(st, q[v], qPZ[v], v == IC ? qMY[v] : qMZ[v]);
    }
  } else {
    for (int d = 0; d < 3; ++d)
      for (int v = 0; v < 8; ++v) dq[d][v] = 0.0;
  }

  real bfNb[6];
  bfNb[0] = U[b.at(i, j, k, IA)];
  bfNb[1] = U[b.at(i + 1, j, k, IA)];
  bfNb[2] = U[b.at(i, j, k, IB)];
  bfNb[3] = U[b.at(i, j + 1, k, IB)];
  bfNb[4] = U[b.at(i, j, k, IC)];
  bfNb[5] = U[b.at(i, j, k + 1, IC)];

  auto D3 = [&](const double *a, int ii, int jj, int kk, int c) {
    return real(a[b.at(ii, jj, kk, c)]);
  };
  real dbf[12];
  dbf[0] = D3(dA, i, j, k, IY);
  dbf[1] = D3(dA, i, j, k, IZ);
  dbf[2] = D3(dB, i, j, k, IX);
  dbf[3] = D3(dB, i, j, k, IZ);
  dbf[4] = D3(dC, i, j, k, IX);
  dbf[5] = D3(dC, i, j, k, IY);
  dbf[6] = D3(dA, i + 1, j, k, IY);
  dbf[7] = D3(dA, i + 1, j, k, IZ);
  dbf[8] = D3(dB, i, j + 1, k, IX);
  dbf[9] = D3(dB, i, j + 1, k, IZ);
  dbf[10] = D3(dC, i, j, k + 1, IX);
  dbf[11] = D3(dC, i, j, k + 1, IY);

  const real ELL = D3(E, i, j, k, IX), ELR = D3(E, i, j, k + 1, IX);
  const real ERL = D3(E, i, j + 1, k, IX), ERR = D3(E, i, j + 1, k + 1, IX);
  const real FLL = D3(E, i, j, k, IY), FLR = D3(E, i, j, k + 1, IY);
  const real FRL = D3(E, i + 1, j, k, IY), FRR = D3(E, i + 1, j, k + 1, IY);
  const real GLL = D3(E, i, j, k, IZ), GLR = D3(E, i, j + 1, k, IZ);
  const real GRL = D3(E, i + 1, j, k, IZ), GRR = D3(E, i + 1, j + 1, k, IZ);

  const real gamma = p.gamma0, smallR = p.smallr, smallp = p.smallp;
  real r = q[ID], pp = q[IP], u = q[IU], v = q[IV], w = q[IW], A = q[IA], B = q[IB], C = q[IC];
  real AL = bfNb[0], AR = bfNb[1], BL = bfNb[2], BR = bfNb[3], CL = bfNb[4], CR = bfNb[5];

  real drx = dq[IX][ID]; drx *= 0.5;
  real dpx = dq[IX][IP]; dpx *= 0.5;
  real dux = dq[IX][IU]; dux *= 0.5;
  real dvx = dq[IX][IV]; dvx *= 0.5;
  real dwx = dq[IX][IW]; dwx *= 0.5;
  real dCx = dq[IX][IC]; dCx *= 0.5;
  real dBx = dq[IX][IB]; dBx *= 0.5;
  real dry = dq[IY][ID]; dry *= 0.5;
  real dpy = dq[IY][IP]; dpy *= 0.5;
  real duy = dq[IY][IU]; duy *= 0.5;
  real dvy = dq[IY][IV]; dvy *= 0.5;
  real dwy = dq[IY][IW]; dwy *= 0.5;
  real dCy = dq[IY][IC]; dCy *= 0.5;
  real dAy = dq[IY][IA]; dAy *= 0.5;
  real drz = dq[IZ][ID]; drz *= 0.5;
  real dpz = dq[IZ][IP]; dpz *= 0.5;
  real duz = dq[IZ][IU]; duz *= 0.5;
  real dvz = dq[IZ][IV]; dvz *= 0.5;
  real dwz = dq[IZ][IW]; dwz *= 0.5;
  real dAz = dq[IZ][IA]; dAz *= 0.5;
  real dBz = dq[IZ][IB]; dBz *= 0.5;

  real dALy = 0.5 * dbf[0], dALz = 0.5 * dbf[1], dBLx = 0.5 * dbf[2], dBLz = 0.5 * dbf[3];
  real dCLx = 0.5 * dbf[4], dCLy = 0.5 * dbf[5];
  real dARy = 0.5 * dbf[6], dARz = 0.5 * dbf[7], dBRx = 0.5 * dbf[8], dBRz = 0.5 * dbf[9];
  real dCRx = 0.5 * dbf[10], dCRy = 0.5 * dbf[11];

  real dAx = 0.5 * (AR - AL);
  real dBy = 0.5 * (BR - BL);
  real dCz = 0.5 * (CR - CL);

  // source terms (:840-879), Omega0 = 0 (never read from the config)
  real sr0 = (-u * drx - dux * r) * dtdx + (-v * dry - dvy * r) * dtdy + (-w * drz - dwz * r) * dtdz;
  real su0 = (-u * dux - (dpx + B * dBx + C * dCx) / r) * dtdx + (-v * duy + B * dAy / r) * dtdy +
             (-w * duz + C * dAz / r) * dtdz;
  real sv0 = (-u * dvx + A * dBx / r) * dtdx + (-v * dvy - (dpy + A * dAy + C * dCy) / r) * dtdy +
             (-w * dvz + C * dBz / r) * dtdz;
  real sw0 = (-u * dwx + A * dCx / r) * dtdx + (-v * dwy + B * dCy / r) * dtdy +
             (-w * dwz - (dpz + A * dAz + B * dBz) / r) * dtdz;
  real sp0 = (-u * dpx - dux * gamma * pp) * dtdx + (-v * dpy - dvy * gamma * pp) * dtdy +
             (-w * dpz - dwz * gamma * pp) * dtdz;
  real sA0 = (u * dBy + B * duy - v * dAy - A * dvy) * dtdy + (u * dCz + C * duz - w * dAz - A * dwz) * dtdz;
  real sB0 = (v * dAx + A * dvx - u * dBx - B * dux) * dtdx + (v * dCz + C * dvz - w * dBz - B * dwz) * dtdz;
  real sC0 = (w * dAx + A * dwx - u * dCx - C * dux) * dtdx + (w * dBy + B * dwy - v * dCy - C * dvy) * dtdy;

  real sAL0 = +(GLR - GLL) * dtdy * 0.5 - (FLR - FLL) * dtdz * 0.5;
  real sAR0 = +(GRR - GRL) * dtdy * 0.5 - (FRR - FRL) * dtdz * 0.5;
  real sBL0 = -(GRL - GLL) * dtdx * 0.5 + (ELR - ELL) * dtdz * 0.5;
  real sBR0 = -(GRR - GLR) * dtdx * 0.5 + (ERR - ERL) * dtdz * 0.5;
  real sCL0 = +(FRL - FLL) * dtdx * 0.5 - (ERL - ELL) * dtdy * 0.5;
  real sCR0 = +(FRR - FLR) * dtdx * 0.5 - (ERR - ELR) * dtdy * 0.5;

  r = r + sr0; u = u + su0; v = v + sv0; w = w + sw0; pp = pp + sp0;
  A = A + sA0; B = B + sB0; C = C + sC0;
  AL = AL + sAL0; AR = AR + sAR0; BL = BL + sBL0; BR = BR + sBR0; CL = CL + sCL0; CR = CR + sCR0;

  // 3D floors: pressure floor is smallp, not smallp*rho (:908 and following)
  auto floor3 = [&](St &s) {
    s[ID] = FMAX(smallR, s[ID]);
    s[IP] = FMAX(smallp, s[IP]);
  };
  St *qp = out.qp, *qm = out.qm;
  qp[0][ID] = r - drx; qp[0][IU] = u - dux; qp[0][IV] = v - dvx; qp[0][IW] = w - dwx;
  qp[0][IP] = pp - dpx; qp[0][IA] = AL; qp[0][IB] = B - dBx; qp[0][IC] = C - dCx;
  floor3(qp[0]);
  qm[0][ID] = r + drx; qm[0][IU] = u + dux; qm[0][IV] = v + dvx; qm[0][IW] = w + dwx;
  qm[0][IP] = pp + dpx; qm[0][IA] = AR; qm[0][IB] = B + dBx; qm[0][IC] = C + dCx;
  floor3(qm[0]);
  qp[1][ID] = r - dry; qp[1][IU] = u - duy; qp[1][IV] = v - dvy; qp[1][IW] = w - dwy;
  qp[1][IP] = pp - dpy; qp[1][IA] = A - dAy; qp[1][IB] = BL; qp[1][IC] = C - dCy;
  floor3(qp[1]);
  qm[1][ID] = r + dry; qm[1][IU] = u + duy; qm[1][IV] = v + dvy; qm[1][IW] = w + dwy;
  qm[1][IP] = pp + dpy; qm[1][IA] = A + dAy; qm[1][IB] = BR; qm[1][IC] = C + dCy;
  floor3(qm[1]);
  qp[2][ID] = r - drz; qp[2][IU] = u - duz; qp[2][IV] = v - dvz; qp[2][IW] = w - dwz;
  qp[2][IP] = pp - dpz; qp[2][IA] = A - dAz; qp[2][IB] = B - dBz; qp[2][IC] = CL;
  floor3(qp[2]);
  qm[2][ID] = r + drz; qm[2][IU] = u + duz; qm[2][IV] = v + dvz; qm[2][IW] = w + dwz;
  qm[2][IP] = pp + dpz; qm[2][IA] = A + dAz; qm[2][IB] = B + dBz; qm[2][IC] = CR;
  floor3(qm[2]);

  // X edges
  {
    St &e = out.qEdge[IRT][0];
    e[ID] = r + (+dry + drz); e[IU] = u + (+duy + duz); e[IV] = v + (+dvy + dvz);
    e[IW] = w + (+dwy + dwz); e[IP] = pp + (+dpy + dpz); e[IA] = A + (+dAy + dAz);
    e[IB] = BR + (+dBRz); e[IC] = CR + (+dCRy);
    floor3(e);
  }
  {
    St &e = out.qEdge[IRB][0];
    e[ID] = r + (+dry - drz); e[IU] = u + (+duy - duz); e[IV] = v + (+dvy - dvz);
    e[IW] = w + (+dwy - dwz); e[IP] = pp + (+dpy - dpz); e[IA] = A + (+dAy - dAz);
    e[IB] = BR + (-dBRz); e[IC] = CL + (+dCLy);
    floor3(e);
  }
  {
    St &e = out.qEdge[ILT][0];
    e[ID] = r + (-dry + drz); e[IU] = u + (-duy + duz); e[IV] = v + (-dvy + dvz);
    e[IW] = w + (-dwy + dwz); e[IP] = pp + (-dpy + dpz); e[IA] = A + (-dAy + dAz);
    e[IB] = BL + (+dBLz); e[IC] = CR + (-dCRy);
    floor3(e);
  }
  {
    St &e = out.qEdge[ILB][0];
    e[ID] = r + (-dry - drz); e[IU] = u + (-duy - duz); e[IV] = v + (-dvy - dvz);
    e[IW] = w + (-dwy - dwz); e[IP] = pp + (-dpy - dpz); e[IA] = A + (-dAy - dAz);
    e[IB] = BL + (-dBLz); e[IC] = CL + (-dCLy);
    floor3(e);
  }
  // Y edges
  {
    St &e = out.qEdge[IRT][1];
    e[ID] = r + (+drx + drz); e[IU] = u + (+dux + duz); e[IV] = v + (+dvx + dvz);
    e[IW] = w + (+dwx + dwz); e[IP] = pp + (+dpx + dpz); e[IA] = AR + (+dARz);
    e[IB] = B + (+dBx + dBz); e[IC] = CR + (+dCRx);
    floor3(e);
  }
  {
    St &e = out.qEdge[IRB][1];
    e[ID] = r + (+drx - drz); e[IU] = u + (+dux - duz); e[IV] = v + (+dvx - dvz);
    e[IW] = w + (+dwx - dwz); e[IP] = pp + (+dpx - dpz); e[IA] = AR + (-dARz);
    e[IB] = B + (+dBx - dBz); e[IC] = CL + (+dCLx);
    floor3(e);
  }
  {
    St &e = out.qEdge[ILT][1];
    e[ID] = r + (-drx + drz); e[IU] = u + (-dux + duz); e[IV] = v + (-dvx + dvz);
    e[IW] = w + (-dwx + dwz); e[IP] = pp + (-dpx + dpz); e[IA] = AL + (+dALz);
    e[IB] = B + (-dBx + dBz); e[IC] = CR + (-dCRx);
    floor3(e);
  }
  {
    St &e = out.qEdge[ILB][1];
    e[ID] = r + (-drx - drz); e[IU] = u + (-dux - duz); e[IV] = v + (-dvx - dvz);
    e[IW] = w + (-dwx - dwz); e[IP] = pp + (-dpx - dpz); e[IA] = AL + (-dALz);
    e[IB] = B + (-dBx - dBz); e[IC] = CL + (-dCLx);
    floor3(e);
  }
  // Z edges
  {
    St &e = out.qEdge[IRT][2];
    e[ID] = r + (+drx + dry); e[IU] = u + (+dux + duy); e[IV] = v + (+dvx + dvy);
    e[IW] = w + (+dwx + dwy); e[IP] = pp + (+dpx + dpy); e[IA] = AR + (+dARy);
    e[IB] = BR + (+dBRx); e[IC] = C + (+dCx + dCy);
    floor3(e);
  }
  {
    St &e = out.qEdge[IRB][2];
    e[ID] = r + (+drx - dry); e[IU] = u + (+dux - duy); e[IV] = v + (+dvx - dvy);
    e[IW] = w + (+dwx - dwy); e[IP] = pp + (+dpx - dpy); e[IA] = AR + (-dARy);
    e[IB] = BL + (+dBLx); e[IC] = C + (+dCx - dCy);
    floor3(e);
  }
  {
    St &e = out.qEdge[ILT][2];
    e[ID] = r + (-drx + dry); e[IU] = u + (-dux + duy); e[IV] = v + (-dvx + dvy);
    e[IW] = w + (-dwx + dwy); e[IP] = pp + (-dpx + dpy); e[IA] = AL + (+dALy);
    e[IB] = BR + (-dBRx); e[IC] = C + (-dCx + dCy);
    floor3(e);
  }
  {
    St &e = out.qEdge[ILB][2];
    e[ID] = r + (-drx - dry); e[IU] = u + (-dux - duy); e[IV] = v + (-dvx - dvy);
    e[IW] = w + (-dwx - dwy); e[IP] = pp + (-dpx - dpy); e[IA] = AL + (-dALy);
    e[IB] = BL + (-dBLx); e[IC] = C + (-dCx - dCy);
    floor3(e);
  }
}

// ---------------------------------------------------------------------------------------------
// solver: SolverMHDMuscl<dim> time loop with impl-0 scratch
// ---------------------------------------------------------------------------------------------
struct Solver {
  ppk_mhd_params p;
  Box b;
  std::vector<double> U, U2, Q;
  std::vector<double> Qm[3], Qp[3], QE[4][3];  // QE[RT..LB][X,Y,Z edge]
  std::vector<double> F[3], Emf, Elec, dA, dB, dC;
  double t = 0.0, dt = 0.0, tEnd = 0.0;
  long long iter = 0;

  explicit Solver(const ppk_mhd_params &pp) : p(pp), b(pp) {
    size_t n8 = (size_t)b.cells() * 8, n3 = (size_t)b.cells() * 3;
    U.assign(n8, 0.0);
    U2.assign(n8, 0.0);
    Q.assign(n8, 0.0);
    int nd = p.dim;
    for (int d = 0; d < nd; ++d) {
      Qm[d].assign(n8, 0.0);
      Qp[d].assign(n8, 0.0);
      F[d].assign(n8, 0.0);
    }
    int ne = nd == 3 ? 3 : 1;
    for (int c = 0; c < 4; ++c)
      for (int d = 0; d < ne; ++d) QE[c][d].assign(n8, 0.0);
    Emf.assign(nd == 3 ? n3 : (size_t)b.cells(), 0.0);
    if (nd == 3) {
      Elec.assign(n3, 0.0);
      dA.assign(n3, 0.0);
      dB.assign(n3, 0.0);
      dC.assign(n3, 0.0);
    }
    tEnd = p.tEnd;
    t = 0.0;
    dt = tEnd;  // SolverBase::read_config, src/shared/SolverBase.cpp:119-121
  }

  // SolverBase::compute_dt (no MPI), src/shared/SolverBase.cpp:149-179
  void compute_dt() {
    double inv = inv_dt(p, Q.data());
    dt = p.cfl / inv;  // src/muscl/SolverMHDMuscl.h:737
    if (t + dt > tEnd) dt = tEnd - t;
  }

  // godunov_unsplit_impl, src/muscl/SolverMHDMuscl.cpp:373-458 (2D) and :465-572 (3D)
  void godunov(double *din, double *dout, double fixed_dt, int dt_mode) {
    ORA_STAGE(0);
    fill_boundaries(p, din);
    std::memcpy(dout, din, sizeof(double) * b.cells() * 8);
    ORA_STAGE(1);
    primitives(p, din, Q.data());
    if (dt_mode == PPK_DT_CFL) {
      compute_dt();
    } else {  // the library clamps a fixed dt to tEnd as SolverBase::compute_dt clamps its own
      dt = fixed_dt;
      if (t + dt > tEnd) dt = tEnd - t;
    }
    const real dtdx = dt / p.dx, dtdy = dt / p.dy, dtdz = p.dim == 3 ? dt / p.dz : 0.0;
    if (p.dim == 2)
      step2d(din, dout, dtdx, dtdy);
    else
      step3d(din, dout, dtdx, dtdy, dtdz);
  }

  void step2d(const double *Uin, double *Uout, real dtdx, real dtdy) {
    const int gw = b.gw, is = b.is, js = b.js;
    ORA_STAGE(4);
    // ComputeTraceFunctor2D_MHD range, src/muscl/MHDRunFunctors2D.h:857-858
#pragma omp parallel for if (!ORA_COUNTING)
    for (int j = gw - 2; j < js - gw + 1; ++j)
      for (int i = gw - 2; i < is - gw + 1; ++i) {
        Trace2D tr;
        trace_cell_2d(p, Uin, Q.data(), i, j, dtdx, dtdy, tr);
        set_state(b, Qm[0].data(), i, j, 0, tr.qm[0]);
        set_state(b, Qp[0].data(), i, j, 0, tr.qp[0]);
        set_state(b, Qm[1].data(), i, j, 0, tr.qm[1]);
        set_state(b, Qp[1].data(), i, j, 0, tr.qp[1]);
        set_state(b, QE[IRT][0].data(), i, j, 0, tr.qEdge[IRT]);
        set_state(b, QE[IRB][0].data(), i, j, 0, tr.qEdge[IRB]);
        set_state(b, QE[ILT][0].data(), i, j, 0, tr.qEdge[ILT]);
        set_state(b, QE[ILB][0].data(), i, j, 0, tr.qEdge[ILB]);
      }
    ORA_STAGE(5);
    // ComputeFluxesAndStoreFunctor2D_MHD, src/muscl/MHDRunFunctors2D.h:420-464
#pragma omp parallel for if (!ORA_COUNTING)
    for (int j = gw; j < js - gw + 1; ++j)
      for (int i = gw; i < is - gw + 1; ++i) {
        St ql, qr, fl;
        get_state(b, Qm[0].data(), i - 1, j, 0, ql);
        get_state(b, Qp[0].data(), i, j, 0, qr);
        riemann_mhd(ql, qr, fl, p);
        set_state(b, F[0].data(), i, j, 0, fl);
        get_state(b, Qm[1].data(), i, j - 1, 0, ql);
        swap_values(ql[IU], ql[IV]);
        swap_values(ql[IA], ql[IB]);
        get_state(b, Qp[1].data(), i, j, 0, qr);
        swap_values(qr[IU], qr[IV]);
        swap_values(qr[IA], qr[IB]);
        riemann_mhd(ql, qr, fl, p);
        set_state(b, F[1].data(), i, j, 0, fl);
      }
    ORA_STAGE(6);
    // ComputeEmfAndStoreFunctor2D, src/muscl/MHDRunFunctors2D.h:654-680
#pragma omp parallel for if (!ORA_COUNTING)
    for (int j = gw; j < js - gw + 1; ++j)
      for (int i = gw; i < is - gw + 1; ++i) {
        St qe[4];
        get_state(b, QE[IRT][0].data(), i - 1, j - 1, 0, qe[IRT]);
        get_state(b, QE[IRB][0].data(), i - 1, j, 0, qe[IRB]);
        get_state(b, QE[ILT][0].data(), i, j - 1, 0, qe[ILT]);
        get_state(b, QE[ILB][0].data(), i, j, 0, qe[ILB]);
        Emf[b.at(i, j, 0, 0)] = D(compute_emf(EMFZ, qe, p));
      }
    ORA_STAGE(7);
    // UpdateFunctor2D_MHD, src/muscl/MHDRunFunctors2D.h:1574-1634
#pragma omp parallel for if (!ORA_COUNTING)
    for (int j = gw; j < js - gw; ++j)
      for (int i = gw; i < is - gw; ++i) {
        St u, f;
        get_state(b, Uout, i, j, 0, u);
        get_state(b, F[0].data(), i, j, 0, f);
        u[ID] += f[ID] * dtdx; u[IP] += f[IP] * dtdx; u[IU] += f[IU] * dtdx;
        u[IV] += f[IV] * dtdx; u[IW] += f[IW] * dtdx; u[IC] += f[IC] * dtdx;
        get_state(b, F[0].data(), i + 1, j, 0, f);
        u[ID] -= f[ID] * dtdx; u[IP] -= f[IP] * dtdx; u[IU] -= f[IU] * dtdx;
        u[IV] -= f[IV] * dtdx; u[IW] -= f[IW] * dtdx; u[IC] -= f[IC] * dtdx;
        get_state(b, F[1].data(), i, j, 0, f);
        u[ID] += f[ID] * dtdy; u[IP] += f[IP] * dtdy; u[IU] += f[IV] * dtdy;
        u[IV] += f[IU] * dtdy; u[IW] += f[IW] * dtdy; u[IC] += f[IC] * dtdy;
        get_state(b, F[1].data(), i, j + 1, 0, f);
        u[ID] -= f[ID] * dtdy; u[IP] -= f[IP] * dtdy; u[IU] -= f[IV] * dtdy;
        u[IV] -= f[IU] * dtdy; u[IW] -= f[IW] * dtdy; u[IC] -= f[IC] * dtdy;
        set_state(b, Uout, i, j, 0, u);
      }
    // UpdateEmfFunctor2D, src/muscl/MHDRunFunctors2D.h:1674-1693
#pragma omp parallel for if (!ORA_COUNTING)
    for (int j = gw; j < js - gw; ++j)
      for (int i = gw; i < is - gw; ++i) {
        real ea = Emf[b.at(i, j, 0, 0)];
        real a = Uout[b.at(i, j, 0, IA)];
        a += (real(Emf[b.at(i, j + 1, 0, 0)]) - ea) * dtdy;
        Uout[b.at(i, j, 0, IA)] = D(a);
        real bb = Uout[b.at(i, j, 0, IB)];
        bb -= (real(Emf[b.at(i + 1, j, 0, 0)]) - ea) * dtdx;
        Uout[b.at(i, j, 0, IB)] = D(bb);
      }
  }

  void step3d(const double *Uin, double *Uout, real dtdx, real dtdy, real dtdz) {
    const int gw = b.gw, is = b.is, js = b.js, ks = b.ks;
    const double *Qd = Q.data();
    ORA_STAGE(2);
    // ComputeElecFieldFunctor3D, src/muscl/MHDRunFunctors3D.h:303-356
#pragma omp parallel for collapse(2) if (!ORA_COUNTING)
    for (int k = 1; k < ks - 1; ++k)
      for (int j = 1; j < js - 1; ++j)
        for (int i = 1; i < is - 1; ++i) {
          auto q = [&](int ii, int jj, int kk, int v) { return real(Qd[b.at(ii, jj, kk, v)]); };
          auto uu = [&](int ii, int jj, int kk, int v) { return real(Uin[b.at(ii, jj, kk, v)]); };
          real u, v, w, A, B, C;
          v = 0.25 * (q(i, j - 1, k - 1, IV) + q(i, j - 1, k, IV) + q(i, j, k - 1, IV) + q(i, j, k, IV));
          w = 0.25 * (q(i, j - 1, k - 1, IW) + q(i, j - 1, k, IW) + q(i, j, k - 1, IW) + q(i, j, k, IW));
          B = 0.5 * (uu(i, j, k - 1, IB) + uu(i, j, k, IB));
          C = 0.5 * (uu(i, j - 1, k, IC) + uu(i, j, k, IC));
          Elec[b.at(i, j, k, IX)] = D(v * C - w * B);
          u = 0.25 * (q(i - 1, j, k - 1, IU) + q(i - 1, j, k, IU) + q(i, j, k - 1, IU) + q(i, j, k, IU));
          w = 0.25 * (q(i - 1, j, k - 1, IW) + q(i - 1, j, k, IW) + q(i, j, k - 1, IW) + q(i, j, k, IW));
          A = 0.5 * (uu(i, j, k - 1, IA) + uu(i, j, k, IA));
          C = 0.5 * (uu(i - 1, j, k, IC) + uu(i, j, k, IC));
          Elec[b.at(i, j, k, IY)] = D(w * A - u * C);
          u = 0.25 * (q(i - 1, j - 1, k, IU) + q(i - 1, j, k, IU) + q(i, j - 1, k, IU) + q(i, j, k, IU));
          v = 0.25 * (q(i - 1, j - 1, k, IV) + q(i - 1, j, k, IV) + q(i, j - 1, k, IV) + q(i, j, k, IV));
          A = 0.5 * (uu(i, j - 1, k, IA) + uu(i, j, k, IA));
          B = 0.5 * (uu(i - 1, j, k, IB) + uu(i, j, k, IB));
          Elec[b.at(i, j, k, IZ)] = D(u * B - v * A);
        }
    ORA_STAGE(3);
    // ComputeMagSlopesFunctor3D + slope_unsplit_mhd_3d, src/muscl/MHDRunFunctors3D.h:472-532,
    // src/muscl/MHDBaseFunctor3D.h:561-668 (xslope_type = min(slope_type, 2))
    const real xst = FMIN(real(p.slope_type), 2.0);
#pragma omp parallel for collapse(2) if (!ORA_COUNTING)
    for (int k = 1; k < ks - 1; ++k)
      for (int j = 1; j < js - 1; ++j)
        for (int i = 1; i < is - 1; ++i) {
          auto uu = [&](int ii, int jj, int kk, int v) { return real(Uin[b.at(ii, jj, kk, v)]); };
          real ax = uu(i, j, k, IA), by = uu(i, j, k, IB), cz = uu(i, j, k, IC);
          real dAy = limited_slope(xst, ax, uu(i, j + 1, k, IA), uu(i, j - 1, k, IA));
          real dAz = limited_slope(xst, ax, uu(i, j, k + 1, IA), uu(i, j, k - 1, IA));
          real dBx = limited_slope(xst, by, uu(i + 1, j, k, IB), uu(i - 1, j, k, IB));
          real dBz = limited_slope(xst, by, uu(i, j, k + 1, IB), uu(i, j, k - 1, IB));
          real dCx = limited_slope(xst, cz, uu(i + 1, j, k, IC), uu(i - 1, j, k, IC));
          real dCy = limited_slope(xst, cz, uu(i, j + 1, k, IC), uu(i, j - 1, k, IC));
          dA[b.at(i, j, k, 0)] = 0.0;
          dA[b.at(i, j, k, 1)] = D(dAy);
          dA[b.at(i, j, k, 2)] = D(dAz);
          dB[b.at(i, j, k, 0)] = D(dBx);
          dB[b.at(i, j, k, 1)] = 0.0;
          dB[b.at(i, j, k, 2)] = D(dBz);
          dC[b.at(i, j, k, 0)] = D(dCx);
          dC[b.at(i, j, k, 1)] = D(dCy);
          dC[b.at(i, j, k, 2)] = 0.0;
        }
    ORA_STAGE(4);
    // ComputeTraceFunctor3D_MHD range, src/muscl/MHDRunFunctors3D.h:678-680
#pragma omp parallel for collapse(2) if (!ORA_COUNTING)
    for (int k = gw - 2; k < ks - gw + 1; ++k)
      for (int j = gw - 2; j < js - gw + 1; ++j)
        for (int i = gw - 2; i < is - gw + 1; ++i) {
          Trace3D tr;
          trace_cell_3d(p, Uin, Qd, Elec.data(), dA.data(), dB.data(), dC.data(), i, j, k,
                        dtdx, dtdy, dtdz, tr);
          for (int d = 0; d < 3; ++d) {
            set_state(b, Qm[d].data(), i, j, k, tr.qm[d]);
            set_state(b, Qp[d].data(), i, j, k, tr.qp[d]);
          }
          for (int c = 0; c < 4; ++c)
            for (int d = 0; d < 3; ++d) set_state(b, QE[c][d].data(), i, j, k, tr.qEdge[c][d]);
        }
    ORA_STAGE(5);
    // ComputeFluxesAndStoreFunctor3D_MHD, src/muscl/MHDRunFunctors3D.h:1839-1902
#pragma omp parallel for collapse(2) if (!ORA_COUNTING)
    for (int k = gw; k < ks - gw + 1; ++k)
      for (int j = gw; j < js - gw + 1; ++j)
        for (int i = gw; i < is - gw + 1; ++i) {
          St ql, qr, fl;
          get_state(b, Qm[0].data(), i - 1, j, k, ql);
          get_state(b, Qp[0].data(), i, j, k, qr);
          riemann_mhd(ql, qr, fl, p);
          set_state(b, F[0].data(), i, j, k, fl);
          get_state(b, Qm[1].data(), i, j - 1, k, ql);
          swap_values(ql[IU], ql[IV]);
          swap_values(ql[IA], ql[IB]);
          get_state(b, Qp[1].data(), i, j, k, qr);
          swap_values(qr[IU], qr[IV]);
          swap_values(qr[IA], qr[IB]);
          riemann_mhd(ql, qr, fl, p);
          set_state(b, F[1].data(), i, j, k, fl);
          get_state(b, Qm[2].data(), i, j, k - 1, ql);
          swap_values(ql[IU], ql[IW]);
          swap_values(ql[IA], ql[IC]);
          get_state(b, Qp[2].data(), i, j, k, qr);
          swap_values(qr[IU], qr[IW]);
          swap_values(qr[IA], qr[IC]);
          riemann_mhd(ql, qr, fl, p);
          set_state(b, F[2].data(), i, j, k, fl);
        }
    ORA_STAGE(6);
    // ComputeEmfAndStoreFunctor3D, src/muscl/MHDRunFunctors3D.h:2183-2230 (EMFY takes RB and
    // LT swapped, :2212-2217)
#pragma omp parallel for collapse(2) if (!ORA_COUNTING)
    for (int k = gw; k < ks - gw + 1; ++k)
      for (int j = gw; j < js - gw + 1; ++j)
        for (int i = gw; i < is - gw + 1; ++i) {
          St qe[4];
          get_state(b, QE[IRT][2].data(), i - 1, j - 1, k, qe[IRT]);
          get_state(b, QE[IRB][2].data(), i - 1, j, k, qe[IRB]);
          get_state(b, QE[ILT][2].data(), i, j - 1, k, qe[ILT]);
          get_state(b, QE[ILB][2].data(), i, j, k, qe[ILB]);
          Emf[b.at(i, j, k, I_EMFZ)] = D(compute_emf(EMFZ, qe, p));
          get_state(b, QE[IRT][1].data(), i - 1, j, k - 1, qe[IRT]);
          get_state(b, QE[ILT][1].data(), i, j, k - 1, qe[IRB]);
          get_state(b, QE[IRB][1].data(), i - 1, j, k, qe[ILT]);
          get_state(b, QE[ILB][1].data(), i, j, k, qe[ILB]);
          Emf[b.at(i, j, k, I_EMFY)] = D(compute_emf(EMFY, qe, p));
          get_state(b, QE[IRT][0].data(), i, j - 1, k - 1, qe[IRT]);
          get_state(b, QE[IRB][0].data(), i, j - 1, k, qe[IRB]);
          get_state(b, QE[ILT][0].data(), i, j, k - 1, qe[ILT]);
          get_state(b, QE[ILB][0].data(), i, j, k, qe[ILB]);
          Emf[b.at(i, j, k, I_EMFX)] = D(compute_emf(EMFX, qe, p));
        }
    ORA_STAGE(7);
    // UpdateFunctor3D_MHD, src/muscl/MHDRunFunctors3D.h:2471-2538
#pragma omp parallel for collapse(2) if (!ORA_COUNTING)
    for (int k = gw; k < ks - gw; ++k)
      for (int j = gw; j < js - gw; ++j)
        for (int i = gw; i < is - gw; ++i) {
          St u, f;
          get_state(b, Uout, i, j, k, u);
          get_state(b, F[0].data(), i, j, k, f);
          u[ID] += f[ID] * dtdx; u[IP] += f[IP] * dtdx; u[IU] += f[IU] * dtdx;
          u[IV] += f[IV] * dtdx; u[IW] += f[IW] * dtdx;
          get_state(b, F[0].data(), i + 1, j, k, f);
          u[ID] -= f[ID] * dtdx; u[IP] -= f[IP] * dtdx; u[IU] -= f[IU] * dtdx;
          u[IV] -= f[IV] * dtdx; u[IW] -= f[IW] * dtdx;
          get_state(b, F[1].data(), i, j, k, f);
          u[ID] += f[ID] * dtdy; u[IP] += f[IP] * dtdy; u[IU] += f[IV] * dtdy;
          u[IV] += f[IU] * dtdy; u[IW] += f[IW] * dtdy;
          get_state(b, F[1].data(), i, j + 1, k, f);
          u[ID] -= f[ID] * dtdy; u[IP] -= f[IP] * dtdy; u[IU] -= f[IV] * dtdy;
          u[IV] -= f[IU] * dtdy; u[IW] -= f[IW] * dtdy;
          get_state(b, F[2].data(), i, j, k, f);
          u[ID] += f[ID] * dtdz; u[IP] += f[IP] * dtdz; u[IU] += f[IW] * dtdz;
          u[IV] += f[IV] * dtdz; u[IW] += f[IU] * dtdz;
          get_state(b, F[2].data(), i, j, k + 1, f);
          u[ID] -= f[ID] * dtdz; u[IP] -= f[IP] * dtdz; u[IU] -= f[IW] * dtdz;
          u[IV] -= f[IV] * dtdz; u[IW] -= f[IU] * dtdz;
          set_state(b, Uout, i, j, k, u);
        }
    // UpdateEmfFunctor3D (CT), src/muscl/MHDRunFunctors3D.h:2583-2621
#pragma omp parallel for collapse(2) if (!ORA_COUNTING)
    for (int k = gw; k < ks - gw; ++k)
      for (int j = gw; j < js - gw; ++j)
        for (int i = gw; i < is - gw; ++i) {
          auto em = [&](int ii, int jj, int kk, int c) { return real(Emf[b.at(ii, jj, kk, c)]); };
          real a = Uout[b.at(i, j, k, IA)], bb = Uout[b.at(i, j, k, IB)], c = Uout[b.at(i, j, k, IC)];
          a += (em(i, j + 1, k, I_EMFZ) - em(i, j, k, I_EMFZ)) * dtdy;
          bb -= (em(i + 1, j, k, I_EMFZ) - em(i, j, k, I_EMFZ)) * dtdx;
          a -= (em(i, j, k + 1, I_EMFY) - em(i, j, k, I_EMFY)) * dtdz;
          bb += (em(i, j, k + 1, I_EMFX) - em(i, j, k, I_EMFX)) * dtdz;
          c += (em(i + 1, j, k, I_EMFY) - em(i, j, k, I_EMFY)) * dtdx;
          c -= (em(i, j + 1, k, I_EMFX) - em(i, j, k, I_EMFX)) * dtdy;
          Uout[b.at(i, j, k, IA)] = D(a);
          Uout[b.at(i, j, k, IB)] = D(bb);
          Uout[b.at(i, j, k, IC)] = D(c);
        }
  }

  double *cur() { return iter % 2 == 0 ? U.data() : U2.data(); }
  const double *cur() const { return iter % 2 == 0 ? U.data() : U2.data(); }

  // SolverBase::next_iteration -> godunov_unsplit ping-pong,
  // src/shared/SolverBase.cpp:206-220, src/muscl/SolverMHDMuscl.h:791-805
  double step(double fixed_dt, int dt_mode) {
    if (iter % 2 == 0)
      godunov(U.data(), U2.data(), fixed_dt, dt_mode);
    else
      godunov(U2.data(), U.data(), fixed_dt, dt_mode);
    ++iter;
    t += dt;
    return dt;
  }
};

// ---------------------------------------------------------------------------------------------
// initial conditions, src/muscl/MHDInitFunctors{2D,3D}.h
// ---------------------------------------------------------------------------------------------
// counter-based generator for the random KH layer: splitmix64 of (seed, global cell, draw).
// The reference draws from a Kokkos XorShift64 pool whose stream depends on the thread
// mapping, so no generator reproduces it; this one is identical on every device.
inline double kh_uniform(uint64_t seed, uint64_t cell, uint64_t draw) {
  uint64_t x = seed * 0x9E3779B97F4A7C15ull + cell * 0xD1B54A32D192ED03ull + draw * 0x8CB92BA72F3D8DD7ull;
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  x = x ^ (x >> 31);
  return (double)(x >> 11) * (1.0 / 9007199254740992.0);
}

void init_problem_array(const ppk_mhd_params &p, const ppk_mhd_problem &pr, double *U) {
  Box b(p);
  const int gw = b.gw;
  std::fill(U, U + b.cells() * 8, 0.0);  // Kokkos views are zero-initialized
  const double gamma0 = p.gamma0;
  const double PI = 3.141592653589793238462643383279502884;
  const double TWOPI = 2 * PI;
  auto at = [&](int i, int j, int k, int v) -> double & { return U[b.at(i, j, k, v)]; };
  const int kend = b.ks;

  if (pr.kind == PPK_PROBLEM_ORSZAG_TANG) {
    if (p.dim == 2) {
      // InitOrszagTangFunctor2D, src/muscl/MHDInitFunctors2D.h:289-383
      const double B0 = 1.0 / std::sqrt(2 * TWOPI) * pr.ot_b0_scale;
      const double p0 = gamma0 / (2 * TWOPI);
      const double d0 = gamma0 * p0;
      const double v0 = 1.0;
      for (int j = 0; j < b.js; ++j)
        for (int i = 0; i < b.is; ++i) {
          double xPos = p.xmin + p.dx / 2 + (i - gw) * p.dx;
          double yPos = p.ymin + p.dy / 2 + (j - gw) * p.dy;
          at(i, j, 0, ID) = d0;
          at(i, j, 0, IU) = -d0 * v0 * std::sin(yPos * TWOPI);
          at(i, j, 0, IV) = d0 * v0 * std::sin(xPos * TWOPI);
          at(i, j, 0, IW) = 0.0;
          at(i, j, 0, IA) = -B0 * std::sin(yPos * TWOPI);
          at(i, j, 0, IB) = B0 * std::sin(2.0 * xPos * TWOPI);
          at(i, j, 0, IC) = 0.0;
        }
      const double TwoPi = 4.0 * std::asin(1.0);
      const double p0e = gamma0 / (2.0 * TwoPi);
      for (int j = 0; j < b.js - 1; ++j)
        for (int i = 0; i < b.is - 1; ++i) {
          double d = at(i, j, 0, ID);
          at(i, j, 0, IP) = p0e / (gamma0 - 1.0) +
                            0.5 * (at(i, j, 0, IU) * at(i, j, 0, IU) / d +
                                   at(i, j, 0, IV) * at(i, j, 0, IV) / d +
                                   0.25 * ((at(i, j, 0, IA) + at(i + 1, j, 0, IA)) *
                                           (at(i, j, 0, IA) + at(i + 1, j, 0, IA))) +
                                   0.25 * ((at(i, j, 0, IB) + at(i, j + 1, 0, IB)) *
                                           (at(i, j, 0, IB) + at(i, j + 1, 0, IB))));
        }
    } else {
      // InitOrszagTangFunctor3D, src/muscl/MHDInitFunctors3D.h:311-409
      const double B0 = 1.0 / std::sqrt(4 * PI) * pr.ot_b0_scale;
      const double p0 = gamma0 / (4 * PI);
      const double d0 = gamma0 * p0;
      const double v0 = 1.0;
      const double kt = pr.ot_kt;
      for (int k = 0; k < kend; ++k)
        for (int j = 0; j < b.js; ++j)
          for (int i = 0; i < b.is; ++i) {
            double xPos = p.xmin + p.dx / 2 + (i - gw) * p.dx;
            double yPos = p.ymin + p.dy / 2 + (j - gw) * p.dy;
            double zPos = p.zmin + p.dz / 2 + (k - gw) * p.dz;
            at(i, j, k, ID) = d0;
            at(i, j, k, IU) = -d0 * v0 * std::sin(yPos * TWOPI);
            at(i, j, k, IV) = d0 * v0 * std::sin(xPos * TWOPI);
            at(i, j, k, IW) = 0.0;
            double cz = std::cos(2 * TWOPI * kt * (zPos - p.zmin) / (p.zmax - p.zmin));
            at(i, j, k, IA) = -B0 * cz * std::sin(yPos * TWOPI);
            at(i, j, k, IB) = B0 * cz * std::sin(2.0 * xPos * TWOPI);
            at(i, j, k, IC) = 0.0;
          }
      for (int k = 0; k < kend; ++k)
        for (int j = 0; j < b.js - 1; ++j)
          for (int i = 0; i < b.is - 1; ++i) {
            double d = at(i, j, k, ID);
            at(i, j, k, IP) = p0 / (gamma0 - 1.0) +
                              0.5 * (at(i, j, k, IU) * at(i, j, k, IU) / d +
                                     at(i, j, k, IV) * at(i, j, k, IV) / d +
                                     0.25 * ((at(i, j, k, IA) + at(i + 1, j, k, IA)) *
                                             (at(i, j, k, IA) + at(i + 1, j, k, IA))) +
                                     0.25 * ((at(i, j, k, IB) + at(i, j + 1, k, IB)) *
                                             (at(i, j, k, IB) + at(i, j + 1, k, IB))));
          }
    }
  } else if (pr.kind == PPK_PROBLEM_BLAST) {
    // InitBlastFunctor{2D,3D}_MHD, src/muscl/MHDInitFunctors2D.h:164-231,
    // src/muscl/MHDInitFunctors3D.h:175-254 (uniform B is 0.5 in the reference)
    const double r2 = pr.blast_radius * pr.blast_radius;
    for (int k = 0; k < kend; ++k)
      for (int j = 0; j < b.js; ++j)
        for (int i = 0; i < b.is; ++i) {
          double x = p.xmin + p.dx / 2 + (i - gw) * p.dx;
          double y = p.ymin + p.dy / 2 + (j - gw) * p.dy;
          double d2;
          if (p.dim == 3) {
            double z = p.zmin + p.dz / 2 + (k - gw) * p.dz;
            d2 = (x - pr.blast_center_x) * (x - pr.blast_center_x) +
                 (y - pr.blast_center_y) * (y - pr.blast_center_y) +
                 (z - pr.blast_center_z) * (z - pr.blast_center_z);
          } else {
            d2 = (x - pr.blast_center_x) * (x - pr.blast_center_x) +
                 (y - pr.blast_center_y) * (y - pr.blast_center_y);
          }
          bool in = d2 < r2;
          at(i, j, k, ID) = in ? pr.blast_density_in : pr.blast_density_out;
          at(i, j, k, IU) = 0.0;
          at(i, j, k, IV) = 0.0;
          at(i, j, k, IW) = 0.0;
          at(i, j, k, IA) = pr.blast_bx;
          at(i, j, k, IB) = pr.blast_by;
          at(i, j, k, IC) = pr.blast_bz;
          double a = at(i, j, k, IA), bb = at(i, j, k, IB), c = at(i, j, k, IC);
          at(i, j, k, IP) = (in ? pr.blast_pressure_in : pr.blast_pressure_out) / (gamma0 - 1.0) +
                            0.5 * (a * a + bb * bb + c * c);
        }
  } else if (pr.kind == PPK_PROBLEM_KELVIN_HELMHOLTZ) {
    // InitKelvinHelmholtzFunctor{2D,3D}_MHD, src/muscl/MHDInitFunctors2D.h:416-565,
    // src/muscl/MHDInitFunctors3D.h:441-604
    const double d_in = pr.kh_d_in, d_out = pr.kh_d_out, vin = pr.kh_vflow_in, vout = pr.kh_vflow_out;
    const double ampl = pr.kh_amplitude, pressure = pr.kh_pressure;
    for (int k = 0; k < kend; ++k)
      for (int j = 0; j < b.js; ++j)
        for (int i = 0; i < b.is; ++i) {
          double x = p.xmin + p.dx / 2 + (i - gw) * p.dx;
          double y = p.ymin + p.dy / 2 + (j - gw) * p.dy;
          double z = p.dim == 3 ? p.zmin + p.dz / 2 + (k - gw) * p.dz : 0.0;
          uint64_t cell = ((uint64_t)k * b.js + j) * b.is + i;
          if (p.dim == 2) {
            if (pr.kh_p_rand) {
              double yn = (y - p.ymin) / (p.ymax - p.ymin);
              double d, u, v;
              if (yn < 0.25 || yn > 0.75) { d = d_out; u = vout; v = 0.0; }
              else { d = d_in; u = vin; v = 0.0; }
              u += ampl * (kh_uniform(pr.kh_seed, cell, 0) - 0.5);
              v += ampl * (kh_uniform(pr.kh_seed, cell, 1) - 0.5);
              const double bx = 0.5, by = 0.0, bz = 0.0;
              at(i, j, k, ID) = d; at(i, j, k, IU) = d * u; at(i, j, k, IV) = d * v;
              at(i, j, k, IA) = bx; at(i, j, k, IB) = by; at(i, j, k, IC) = bz;
              at(i, j, k, IP) = pressure / (gamma0 - 1.0) + 0.5 * d * (u * u + v * v) +
                                0.5 * (bx * bx + by * by + bz * bz);
            } else if (pr.kh_p_sine_rob || pr.kh_p_sine) {
              const int n = pr.kh_mode;
              const double w0 = pr.kh_w0, delta = pr.kh_delta, y1 = 0.25, y2 = 0.75;
              double d, u;
              if (pr.kh_p_sine_rob) {
                const double ramp = 1.0 / (1.0 + std::exp(2 * (y - y1) / delta)) +
                                    1.0 / (1.0 + std::exp(2 * (y2 - y) / delta));
                d = d_in + ramp * (d_out - d_in);
                u = vin + ramp * (vout - vin);
              } else {
                d = (y >= y1 && y <= y2) ? d_in : d_out;
                u = (y >= y1 && y <= y2) ? vin : vout;
              }
              const double v = w0 * std::sin(n * PI * x);
              const double bx = 0.5, by = 0.0, bz = 0.0;
              at(i, j, k, ID) = d; at(i, j, k, IU) = d * u; at(i, j, k, IV) = d * v;
              at(i, j, k, IA) = bx; at(i, j, k, IB) = by; at(i, j, k, IC) = bz;
              at(i, j, k, IP) = pressure / (gamma0 - 1.0) + 0.5 * d * (u * u + v * v) +
                                0.5 * (bx * bx + by * by + bz * bz);
            }
          } else {
            // shear-normal coordinate: z in the reference; y when kh_axis == 1 (extension)
            const bool yaxis = pr.kh_axis == 1;
            double s = yaxis ? y : z;
            double smin = yaxis ? p.ymin : p.zmin, smax = yaxis ? p.ymax : p.zmax;
            if (pr.kh_p_rand) {
              double sn = (s - smin) / (smax - smin);
              double d, u, v, w;
              if (sn < 0.25 || sn > 0.75) { d = d_out; u = vout; v = 0.0; w = 0.0; }
              else { d = d_in; u = vin; v = 0.0; w = 0.0; }
              u += ampl * (kh_uniform(pr.kh_seed, cell, 0) - 0.5);
              v += ampl * (kh_uniform(pr.kh_seed, cell, 1) - 0.5);
              if (!yaxis) w += ampl * (kh_uniform(pr.kh_seed, cell, 2) - 0.5);
              const double bx = pr.kh_bx, by = pr.kh_by, bz = pr.kh_bz;
              at(i, j, k, ID) = d; at(i, j, k, IU) = d * u; at(i, j, k, IV) = d * v;
              at(i, j, k, IW) = d * w;
              at(i, j, k, IA) = bx; at(i, j, k, IB) = by; at(i, j, k, IC) = bz;
              at(i, j, k, IP) = pressure / (gamma0 - 1.0) + 0.5 * d * (u * u + v * v + w * w) +
                                0.5 * (bx * bx + by * by + bz * bz);
            } else if (pr.kh_p_sine_rob) {
              const int n = pr.kh_mode;
              const double w0 = pr.kh_w0, delta = pr.kh_delta, z1 = 0.25, z2 = 0.75;
              const double ramp = 1.0 / (1.0 + std::exp(2 * (z - z1) / delta)) +
                                  1.0 / (1.0 + std::exp(2 * (z2 - z) / delta));
              const double v1y = vin / 2, v2y = vout / 2;
              const double d = d_in + ramp * (d_out - d_in);
              const double u = vin + ramp * (vout - vin);
              const double v = v1y + ramp * (v2y - v1y);
              const double w = w0 * std::sin(n * M_PI * x) * std::sin(n * M_PI * y);
              const double bx = 0.5, by = 0.0, bz = 0.0;
              at(i, j, k, ID) = d; at(i, j, k, IU) = d * u; at(i, j, k, IV) = d * v;
              at(i, j, k, IW) = d * w;
              at(i, j, k, IA) = bx; at(i, j, k, IB) = by; at(i, j, k, IC) = bz;
              at(i, j, k, IP) = pressure / (gamma0 - 1.0) + 0.5 * d * (u * u + v * v + w * w) +
                                0.5 * (bx * bx + by * by + bz * bz);
            } else if (pr.kh_p_sine) {
              const int n = pr.kh_mode;
              const double w0 = pr.kh_w0, z1 = 0.25, z2 = 0.75;
              const double d = (z >= z1 && z <= z2) ? d_in : d_out;
              const double u = (z >= z1 && z <= z2) ? vin : vout;
              const double v = 0;
              const double w = w0 * std::sin(n * M_PI * x);
              const double bx = 0.5, by = 0.0, bz = 0.0;
              at(i, j, k, ID) = d; at(i, j, k, IU) = d * u; at(i, j, k, IV) = d * v;
              at(i, j, k, IW) = d * w;
              at(i, j, k, IA) = bx; at(i, j, k, IB) = by; at(i, j, k, IC) = bz;
              at(i, j, k, IP) = pressure / (gamma0 - 1.0) + 0.5 * d * (u * u + v * v + w * w) +
                                0.5 * (bx * bx + by * by + bz * bz);
            }
          }
        }
  } else if (pr.kind == PPK_PROBLEM_WAVE) {
    if (p.dim == 2) return;  // InitWaveFunctor2D_MHD does nothing (:972-975)
    // WaveParams, src/shared/problems/WaveParams.h:31-146
    const double wg = pr.wave_gamma0;
    double rev[7];
    double V0;
    switch (pr.wave_type) {
      case 0: { double r[7] = {4.472136e-01, -8.944272e-01, 4.216370e-01, 1.490712e-01,
                               2.012461e+00, 8.432740e-01, 2.981424e-01};
                std::memcpy(rev, r, sizeof r); V0 = 0.0; break; }
      case 1: { double r[7] = {0.0, 0.0, -3.333333e-01, 9.428090e-01, 0.0, -3.333333e-01,
                               9.428090e-01};
                std::memcpy(rev, r, sizeof r); V0 = 0.0; break; }
      case 2: { double r[7] = {8.944272e-01, -4.472136e-01, -8.432740e-01, -2.981424e-01,
                               6.708204e-01, -4.216370e-01, -1.490712e-01};
                std::memcpy(rev, r, sizeof r); V0 = 0.0; break; }
      default: { double r[7] = {1.0, 1.0, 0.0, 0.0, 0.5, 0.0, 0.0};
                 std::memcpy(rev, r, sizeof r); V0 = 1.0; break; }
    }
    const double amp = pr.wave_amplitude;
    const double d0 = 1.0, p0 = 1.0 / wg;
    const double TwoPi = 4.0 * std::asin(1.0);
    const double Lx = pr.wave_lx, Ly = pr.wave_ly, Lz = pr.wave_lz;
    double ang_3 = std::atan(Lx / Ly);
    double sin_a3 = std::sin(ang_3), cos_a3 = std::cos(ang_3);
    double ang_2 = std::atan(0.5 * (Lx * cos_a3 + Ly * sin_a3) / Lz);
    double sin_a2 = std::sin(ang_2), cos_a2 = std::cos(ang_2);
    double x1 = Lx * cos_a2 * cos_a3, x2 = Ly * cos_a2 * sin_a3, x3 = Lz * sin_a2;
    double lambda = x1;
    if (ang_3 != 0.0) lambda = std::fmin(lambda, x2);
    if (ang_2 != 0.0) lambda = std::fmin(lambda, x3);
    const double k_par = TwoPi / lambda;
    const double dby = amp * rev[5], dbz = amp * rev[6];
    const double bx0 = 1.0, by0 = std::sqrt(2.0), bz0 = 0.5;
    // InitWaveFunctor3D_MHD phases, src/muscl/MHDInitFunctors3D.h:1098-1298
    std::vector<double> A((size_t)b.cells() * 3, 0.0);
    auto Aat = [&](int i, int j, int k, int c) -> double & { return A[b.at(i, j, k, c)]; };
    for (int k = 0; k < b.ks; ++k)
      for (int j = 0; j < b.js; ++j)
        for (int i = 0; i < b.is; ++i) {
          double x = p.xmin + p.dx / 2 + (i - gw) * p.dx;
          double y = p.ymin + p.dy / 2 + (j - gw) * p.dy;
          double z = p.zmin + p.dz / 2 + (k - gw) * p.dz;
          double Ay, Az, tmpx, tmpy, X1, X2, X3;
          X1 = x; X2 = y - p.dy / 2; X3 = z - p.dz / 2;
          tmpx = X1 * cos_a2 * cos_a3 + X2 * cos_a2 * sin_a3 + X3 * sin_a2;
          tmpy = -X1 * sin_a3 + X2 * cos_a3;
          Ay = bz0 * tmpx - (dbz / k_par) * std::cos(k_par * tmpx);
          Az = -by0 * tmpx + (dby / k_par) * std::cos(k_par * tmpx) + bx0 * tmpy;
          Aat(i, j, k, IX) = -Ay * sin_a3 - Az * sin_a2 * cos_a3;
          X1 = x - p.dx / 2; X2 = y; X3 = z - p.dz / 2;
          tmpx = X1 * cos_a2 * cos_a3 + X2 * cos_a2 * sin_a3 + X3 * sin_a2;
          tmpy = -X1 * sin_a3 + X2 * cos_a3;
          Ay = bz0 * tmpx - (dbz / k_par) * std::cos(k_par * tmpx);
          Az = -by0 * tmpx + (dby / k_par) * std::cos(k_par * tmpx) + bx0 * tmpy;
          Aat(i, j, k, IY) = Ay * cos_a3 - Az * sin_a2 * sin_a3;
          X1 = x - p.dx / 2; X2 = y - p.dy / 2; X3 = z;
          tmpx = X1 * cos_a2 * cos_a3 + X2 * cos_a2 * sin_a3 + X3 * sin_a2;
          tmpy = -X1 * sin_a3 + X2 * cos_a3;
          Az = -by0 * tmpx + (dby / k_par) * std::cos(k_par * tmpx) + bx0 * tmpy;
          Aat(i, j, k, IZ) = Az * cos_a2;
        }
    for (int k = gw - 1; k < b.ks - gw + 1; ++k)
      for (int j = gw - 1; j < b.js - gw + 1; ++j)
        for (int i = gw - 1; i < b.is - gw + 1; ++i) {
          at(i, j, k, IA) = (Aat(i, j + 1, k, 2) - Aat(i, j, k, 2)) / p.dy -
                            (Aat(i, j, k + 1, 1) - Aat(i, j, k, 1)) / p.dz;
          at(i, j, k, IB) = (Aat(i, j, k + 1, 0) - Aat(i, j, k, 0)) / p.dz -
                            (Aat(i + 1, j, k, 2) - Aat(i, j, k, 2)) / p.dx;
          at(i, j, k, IC) = (Aat(i + 1, j, k, 1) - Aat(i, j, k, 1)) / p.dx -
                            (Aat(i, j + 1, k, 0) - Aat(i, j, k, 0)) / p.dy;
        }
    for (int k = gw; k < b.ks - gw; ++k)
      for (int j = gw; j < b.js - gw; ++j)
        for (int i = gw; i < b.is - gw; ++i) {
          double x = p.xmin + p.dx / 2 + (i - gw) * p.dx;
          double y = p.ymin + p.dy / 2 + (j - gw) * p.dy;
          double z = p.zmin + p.dz / 2 + (k - gw) * p.dz;
          double X = cos_a2 * (x * cos_a3 + y * sin_a3) + z * sin_a2;
          double sn = std::sin(k_par * X);
          double Mx = d0 * V0 + amp * sn * rev[1];
          double My = amp * sn * rev[2];
          double Mz = amp * sn * rev[3];
          at(i, j, k, ID) = d0 + amp * sn * rev[0];
          at(i, j, k, IU) = Mx * cos_a2 * cos_a3 - My * sin_a3 - Mz * sin_a2 * cos_a3;
          at(i, j, k, IV) = Mx * cos_a2 * sin_a3 + My * cos_a3 - Mz * sin_a2 * sin_a3;
          at(i, j, k, IW) = Mx * sin_a2 + Mz * cos_a2;
          at(i, j, k, IP) = p0 / (gamma0 - 1.0) + 0.5 * d0 * V0 * V0 +
                            0.5 * (bx0 * bx0 + by0 * by0 + bz0 * bz0) + amp * sn * rev[4];
        }
  } else if (pr.kind == PPK_PROBLEM_IMPLODE) {
    // InitImplodeFunctor{2D,3D}_MHD, src/muscl/MHDInitFunctors2D.h:58-134,
    // src/muscl/MHDInitFunctors3D.h:56-145
    const double *in = pr.implode_in, *out = pr.implode_out;
    for (int k = 0; k < kend; ++k)
      for (int j = 0; j < b.js; ++j)
        for (int i = 0; i < b.is; ++i) {
          double x = p.xmin + p.dx / 2 + (i - gw) * p.dx;
          double y = p.ymin + p.dy / 2 + (j - gw) * p.dy;
          bool tmp;
          if (p.dim == 2) {
            if (pr.implode_shape == 1) tmp = x + y > 0.5 && x + y < 2.5;
            else tmp = x + y > (p.xmin + p.xmax) / 2. + p.ymin;
          } else {
            double z = p.zmin + p.dz / 2 + (k - gw) * p.dz;
            if (pr.implode_shape == 1) tmp = x + y + z > 0.5 && x + y + z < 2.5;
            else tmp = x + y + z > (p.xmin + p.xmax) / 2. + p.ymin + p.zmin;
          }
          const double *s = tmp ? out : in;
          at(i, j, k, ID) = s[0];
          if (p.dim == 2)
            at(i, j, k, IP) = s[1] / (gamma0 - 1.0) + 0.5 * s[0] * (s[2] * s[2] + s[3] * s[3]) +
                              0.5 * (s[5] * s[5] + s[6] * s[6]);
          else
            at(i, j, k, IP) = s[1] / (gamma0 - 1.0) +
                              0.5 * s[0] * (s[2] * s[2] + s[3] * s[3] + s[4] * s[4]) +
                              0.5 * (s[5] * s[5] + s[6] * s[6] + s[7] * s[7]);
          at(i, j, k, IU) = s[2];
          at(i, j, k, IV) = s[3];
          at(i, j, k, IW) = p.dim == 2 ? 0.0 : s[4];
          at(i, j, k, IA) = s[5];
          at(i, j, k, IB) = s[6];
          at(i, j, k, IC) = p.dim == 2 ? 0.0 : s[7];
        }
  } else if (pr.kind == PPK_PROBLEM_ROTOR) {
    // InitRotorFunctor2D_MHD (src/muscl/MHDInitFunctors2D.h:600-712) and InitRotorFunctor3D_MHD
    // (src/muscl/MHDInitFunctors3D.h:640-756)
    const double r0 = pr.rotor_r0, r1 = pr.rotor_r1, u0 = pr.rotor_u0, p0 = pr.rotor_p0;
    const double xCenter = (p.xmax + p.xmin) / 2, yCenter = (p.ymax + p.ymin) / 2;
    for (int k = 0; k < kend; ++k)
      for (int j = 0; j < b.js; ++j)
        for (int i = 0; i < b.is; ++i) {
          const double x = p.xmin + p.dx / 2 + (i - gw) * p.dx;
          const double y = p.ymin + p.dy / 2 + (j - gw) * p.dy;
          const double r = std::sqrt((x - xCenter) * (x - xCenter) + (y - yCenter) * (y - yCenter));
          const double f_r = (r1 - r) / (r1 - r0);
          if (p.dim == 2) {
            if (r <= r0) {
              at(i, j, k, ID) = 10.0;
              at(i, j, k, IU) = -at(i, j, k, ID) * f_r * u0 * (y - yCenter) / r0;
              at(i, j, k, IV) = at(i, j, k, ID) * f_r * u0 * (x - xCenter) / r0;
            } else if (r <= r1) {
              at(i, j, k, ID) = 1 + 9 * f_r;
              at(i, j, k, IU) = -at(i, j, k, ID) * f_r * u0 * (y - yCenter) / r;
              at(i, j, k, IV) = at(i, j, k, ID) * f_r * u0 * (x - xCenter) / r;
            } else {
              at(i, j, k, ID) = 1.0;
              at(i, j, k, IU) = 0.0;
              at(i, j, k, IV) = 0.0;
            }
          } else {
            if (r <= r0) {
              at(i, j, k, ID) = 10.0;
              at(i, j, k, IU) = -u0 * (y - yCenter) / r0;
              at(i, j, k, IV) = u0 * (x - xCenter) / r0;
            } else if (r <= r1) {
              at(i, j, k, ID) = 1 + 9 * f_r;
              at(i, j, k, IU) = -f_r * u0 * (y - yCenter) / r;
              at(i, j, k, IV) = f_r * u0 * (x - xCenter) / r;
            } else {
              at(i, j, k, ID) = 1.0;
              at(i, j, k, IU) = 0.0;
              at(i, j, k, IV) = 0.0;
            }
          }
          at(i, j, k, IW) = 0.0;
          at(i, j, k, IA) = pr.rotor_b0;
          at(i, j, k, IB) = 0.0;
          at(i, j, k, IC) = 0.0;
          const double U = at(i, j, k, IU), V = at(i, j, k, IV), W = at(i, j, k, IW);
          const double Ab = at(i, j, k, IA);
          if (p.dim == 2)
            at(i, j, k, IP) = p0 / (gamma0 - 1.0) + 0.5 * (U * U + V * V + W * W) / at(i, j, k, ID) +
                              0.5 * (Ab * Ab);
          else
            at(i, j, k, IP) = p0 / (gamma0 - 1.0) + (U * U + V * V + W * W) / 2 / at(i, j, k, ID) +
                              (Ab * Ab) / 2;
        }
  } else if (pr.kind == PPK_PROBLEM_FIELD_LOOP) {
    // InitFieldLoopFunctor2D_MHD (src/muscl/MHDInitFunctors2D.h:715-893) and
    // InitFieldLoopFunctor3D_MHD (src/muscl/MHDInitFunctors3D.h:759-1010): vector potential on
    // every cell, then the interior state with the curl of A, then the interior energy
    std::vector<double> A((size_t)b.cells() * 3, 0.0);
    auto Aat = [&](int i, int j, int k, int c) -> double & { return A[b.at(i, j, k, c)]; };
    for (int k = 0; k < kend; ++k)
      for (int j = 0; j < b.js; ++j)
        for (int i = 0; i < b.is; ++i) {
          const double x = p.xmin + p.dx / 2 + (i - gw) * p.dx;
          const double y = p.ymin + p.dy / 2 + (j - gw) * p.dy;
          const double r = std::sqrt(x * x + y * y);
          if (r < pr.loop_radius) Aat(i, j, k, 2) = pr.loop_amplitude * (pr.loop_radius - r);
        }
    const double cos_theta = 2.0 / std::sqrt(5.0);
    const double sin_theta = std::sqrt(1 - cos_theta * cos_theta);
    const int k0 = p.dim == 3 ? gw : 0, k1 = p.dim == 3 ? b.ks - gw : 1;
    for (int k = k0; k < k1; ++k)
      for (int j = gw; j < b.js - gw; ++j)
        for (int i = gw; i < b.is - gw; ++i) {
          const double x = p.xmin + p.dx / 2 + (i - gw) * p.dx;
          const double y = p.ymin + p.dy / 2 + (j - gw) * p.dy;
          const double r = std::sqrt(x * x + y * y);
          at(i, j, k, ID) = r < pr.loop_radius ? pr.loop_density_in : 1.0;
          at(i, j, k, IU) = at(i, j, k, ID) * pr.loop_vflow * cos_theta;
          at(i, j, k, IV) = at(i, j, k, ID) * pr.loop_vflow * sin_theta;
          if (p.dim == 2) {
            const double diag = std::sqrt(1.0 * (p.nx * p.nx + p.ny * p.ny + p.nz * p.nz));
            at(i, j, k, IW) = at(i, j, k, ID) * pr.loop_vflow * p.nz / diag;
            at(i, j, k, IA) = (Aat(i, j + 1, k, 2) - Aat(i, j, k, 2)) / p.dy;
            at(i, j, k, IB) = -(Aat(i + 1, j, k, 2) - Aat(i, j, k, 2)) / p.dx;
            at(i, j, k, IC) = 0.0;
          } else {
            at(i, j, k, IW) = at(i, j, k, ID) * pr.loop_vflow;
            at(i, j, k, IA) = (Aat(i, j + 1, k, 2) - Aat(i, j, k, 2)) / p.dy -
                              (Aat(i, j, k + 1, 1) - Aat(i, j, k, 1)) / p.dz;
            at(i, j, k, IB) = (Aat(i, j, k + 1, 0) - Aat(i, j, k, 0)) / p.dz -
                              (Aat(i + 1, j, k, 2) - Aat(i, j, k, 2)) / p.dx;
            at(i, j, k, IC) = (Aat(i + 1, j, k, 1) - Aat(i, j, k, 1)) / p.dx -
                              (Aat(i, j + 1, k, 0) - Aat(i, j, k, 0)) / p.dy;
          }
        }
    for (int k = k0; k < k1; ++k)
      for (int j = gw; j < b.js - gw; ++j)
        for (int i = gw; i < b.is - gw; ++i) {
          if (p.cIso > 0) {
            at(i, j, k, IP) = 0.0;
            continue;
          }
          auto sqr = [](double v) { return v * v; };
          const double U = at(i, j, k, IU), V = at(i, j, k, IV), W = at(i, j, k, IW);
          if (p.dim == 2)
            at(i, j, k, IP) = 1.0f / (gamma0 - 1.0) +
                              0.5 * (0.25 * sqr(at(i, j, k, IA) + at(i + 1, j, k, IA)) +
                                     0.25 * sqr(at(i, j, k, IB) + at(i, j + 1, k, IB))) +
                              0.5 * (U * U + V * V) / at(i, j, k, ID);
          else
            at(i, j, k, IP) = 1.0f / (gamma0 - 1.0) +
                              0.5 * (0.25 * sqr(at(i, j, k, IA) + at(i + 1, j, k, IA)) +
                                     0.25 * sqr(at(i, j, k, IB) + at(i, j + 1, k, IB)) +
                                     0.25 * sqr(at(i, j, k, IC) + at(i, j, k + 1, IC))) +
                              0.5 * (U * U + V * V + W * W) / at(i, j, k, ID);
        }
  }
}

St load8(const double *q) {
  St s;
  for (int v = 0; v < 8; ++v) s[v] = q[v];
  return s;
}

}  // namespace

struct ora_solver {
  Solver s;
  explicit ora_solver(const ppk_mhd_params &p) : s(p) {}
};

extern "C" {

// OpenMP threads the oracle's parallel loops use (the CPU baseline records it)
int ora_threads() {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

int ora_params_size(void) { return (int)sizeof(ppk_mhd_params); }
int ora_problem_size(void) { return (int)sizeof(ppk_mhd_problem); }

void ora_fill_boundaries(const ppk_mhd_params *p, double *U) { fill_boundaries(*p, U); }
void ora_primitives(const ppk_mhd_params *p, const double *U, double *Q) { primitives(*p, U, Q); }
double ora_inv_dt(const ppk_mhd_params *p, const double *Q) { return inv_dt(*p, Q); }
void ora_init_problem_array(const ppk_mhd_params *p, const ppk_mhd_problem *prob, double *U) {
  init_problem_array(*p, *prob, U);
}

void ora_riemann(const ppk_mhd_params *p, const double *qleft, const double *qright,
                 double *flux) {
  St l = load8(qleft), r = load8(qright), f;
  riemann_mhd(l, r, f, *p);
  for (int v = 0; v < 8; ++v) flux[v] = D(f[v]);
}

double ora_compute_emf(const ppk_mhd_params *p, int emf_dir, const double *qedge) {
  St q[4];
  for (int c = 0; c < 4; ++c) q[c] = load8(qedge + 8 * c);
  return D(compute_emf(emf_dir, q, *p));
}

double ora_mag_riemann2d_hlld(const ppk_mhd_params *p, const double *qLLRR,
                              const double *eLLRR) {
  St q[4];
  real e[4];
  for (int c = 0; c < 4; ++c) {
    q[c] = load8(qLLRR + 8 * c);
    e[c] = eLLRR[c];
  }
  return D(mag_riemann2d_hlld(q, e, *p));
}

double ora_find_speed_fast(const ppk_mhd_params *p, const double *q, int dir) {
  return D(speed_fast(load8(q), dir, p->gamma0));
}

double ora_slope(const ppk_mhd_params *p, double q, double qplus, double qminus) {
  return D(limited_slope(p->slope_type, q, qplus, qminus));
}

ora_solver *ora_create(const ppk_mhd_params *p) { return new ora_solver(*p); }
void ora_destroy(ora_solver *s) { delete s; }

// SolverMHDMuscl ctor sequence after allocation: init, boundaries, U2 = U, primitives,
// compute_dt, src/muscl/SolverMHDMuscl.h:389-402
void ora_init_problem(ora_solver *s, const ppk_mhd_problem *prob) {
  Solver &S = s->s;
  init_problem_array(S.p, *prob, S.U.data());
  fill_boundaries(S.p, S.U.data());
  S.U2 = S.U;
  S.t = 0.0;
  S.iter = 0;
  primitives(S.p, S.U.data(), S.Q.data());
  S.compute_dt();
}

void ora_set_state(ora_solver *s, const double *U) {
  Solver &S = s->s;
  std::memcpy(S.cur(), U, sizeof(double) * S.b.cells() * 8);
}
void ora_get_state(const ora_solver *s, double *U) {
  const Solver &S = s->s;
  std::memcpy(U, S.cur(), sizeof(double) * S.b.cells() * 8);
}
double ora_step(ora_solver *s, double dt, int dt_mode) { return s->s.step(dt, dt_mode); }
void ora_get_time(const ora_solver *s, double *t, double *dt, long long *iteration) {
  if (t) *t = s->s.t;
  if (dt) *dt = s->s.dt;
  if (iteration) *iteration = s->s.iter;
}
void ora_set_time(ora_solver *s, double t, long long iteration) {
  s->s.t = t;
  s->s.iter = iteration;
}
// SolverBase::finished, src/shared/SolverBase.cpp:195-201
int ora_finished(const ora_solver *s) {
  return s->s.t >= (s->s.tEnd - 1e-14) || s->s.iter >= s->s.p.n_step_max;
}
const double *ora_scratch(const ora_solver *s, int which) {
  const Solver &S = s->s;
  switch (which) {
    case 0: return S.Q.data();
    case 1: return S.F[0].data();
    case 2: return S.F[1].data();
    case 3: return S.p.dim == 3 ? S.F[2].data() : nullptr;
    case 4: return S.Emf.data();
    default: return nullptr;
  }
}

void ora_op_counts(long long counts[5]) {
  ora_op_counts_stage(-1, counts);
}
void ora_op_counts_stage(int stage, long long counts[5]) {
  for (int c = 0; c < 5; ++c) counts[c] = 0;
#ifdef ORA_COUNT
  for (int st = 0; st < kStages; ++st) {
    if (stage >= 0 && st != stage) continue;
    counts[0] += g_stage_ops[st].add;
    counts[1] += g_stage_ops[st].mul;
    counts[2] += g_stage_ops[st].div;
    counts[3] += g_stage_ops[st].sqrt;
    counts[4] += g_stage_ops[st].sel;
  }
#else
  (void)stage;
#endif
}
void ora_op_counts_reset(void) {
#ifdef ORA_COUNT
  for (int st = 0; st < kStages; ++st) g_stage_ops[st] = OpCounts();
  g_stage = 0;
#endif
}
int ora_is_counting_build(void) {
#ifdef ORA_COUNT
  return 1;
#else
  return 0;
#endif
}

}  // extern "C"
