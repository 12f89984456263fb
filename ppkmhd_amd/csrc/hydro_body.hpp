// Hydro (Euler) MUSCL-Hancock building blocks, host+device: primitive variables, TVD slopes,
// the trace of a cell to one of its faces, and the Riemann solvers (approximate two-shock,
// LLF, HLL, HLLC).  Same operations in the same order as the reference
// (src/muscl/HydroBaseFunctor{2D,3D}.h, src/shared/RiemannSolvers.h), so a build with
// -ffp-contract=off and PPK_EXACT_DIV reproduces the reference's double results bit for bit.
// Quotients and square roots go through divr / recip / msqrt (mhd_device.hpp): exactly a / b
// and sqrt in that build; in the default build a denominator used several times is inverted
// once and the GPU uses refined reciprocal / reciprocal-square-root sequences.
// NV = 4 (2D: ID, IP, IU, IV) or 5 (3D: + IW).
#pragma once

#include <cmath>

#include "mhd_device.hpp"

namespace ppk {
namespace hyd {

template <int NV>
struct HSt {
  double v[NV];
  PPK_HD double &operator[](int i) { return v[i]; }
  PPK_HD const double &operator[](int i) const { return v[i]; }
};

enum { FACE_XMIN = 0, FACE_XMAX = 1, FACE_YMIN = 2, FACE_YMAX = 3, FACE_ZMIN = 4, FACE_ZMAX = 5 };

// the parameters the hydro kernels read
struct HydroPar {
  double gamma0, gamma6, smallr, smallc, smallp, smallpp, slope_type;
  double entho;  // 1 / (gamma0 - 1), as each solver computes it (same IEEE value)
  int riemann;  // enum ppk_riemann
};

// computePrimitives (HydroBaseFunctor2D.h:81-105, HydroBaseFunctor3D.h:81-106)
template <int NV>
PPK_HD void prims(const HSt<NV> &u, const HydroPar &P, double &c, HSt<NV> &q) {
  const double d = fmax(u[ID], P.smallr);
  const double id = recip(d);
  const double ux = divr(u[IU], d, id);
  const double uy = divr(u[IV], d, id);
  double eken;
  double uz = 0.0;
  if constexpr (NV == 5) {
    uz = divr(u[IW], d, id);
    eken = 0.5 * (ux * ux + uy * uy + uz * uz);
  } else {
    eken = 0.5 * (ux * ux + uy * uy);
  }
  const double e = divr(u[IP], d, id) - eken;
  const double p = fmax((P.gamma0 - 1.0) * d * e, d * P.smallp);
  c = msqrt(divr(P.gamma0 * (p), d, id));
  q[ID] = d;
  q[IP] = p;
  q[IU] = ux;
  q[IV] = uy;
  if constexpr (NV == 5) q[IW] = uz;
}

// one direction of slope_unsplit_hydro_{2d,3d}_scalar (HydroBaseFunctor2D.h:423-459)
PPK_HD double slope1(double q, double qp, double qm, double slope_type) {
  const double dlft = slope_type * (q - qm);
  const double drgt = slope_type * (qp - q);
  const double dcen = 0.5 * (qp - qm);
  const double dsgn = (dcen >= 0.0) ? 1.0 : -1.0;
  const double slop = fmin(fabs(dlft), fabs(drgt));
  double dlim = slop;
  if ((dlft * drgt) <= 0.0) dlim = 0.0;
  return dsgn * fmin(dlim, fabs(dcen));
}

// the slopes of a cell from its neighbours (slope_type 0: none; 1, 2: minmod / average)
template <int NV>
PPK_HD void slopes(const HSt<NV> &q, const HSt<NV> &qpx, const HSt<NV> &qmx, const HSt<NV> &qpy,
                   const HSt<NV> &qmy, const HSt<NV> &qpz, const HSt<NV> &qmz,
                   const HydroPar &P, HSt<NV> &dx, HSt<NV> &dy, HSt<NV> &dz) {
  // NV == 4 (2D): dz is not touched
  const bool flat = P.slope_type == 0;
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    dx[v] = flat ? 0.0 : slope1(q[v], qpx[v], qmx[v], P.slope_type);
    dy[v] = flat ? 0.0 : slope1(q[v], qpy[v], qmy[v], P.slope_type);
    if constexpr (NV == 5) dz[v] = flat ? 0.0 : slope1(q[v], qpz[v], qmz[v], P.slope_type);
  }
}

// Hancock source terms of trace_unsplit_{2d,3d}_along_dir (HydroBaseFunctor2D.h:218-301,
// HydroBaseFunctor3D.h:276-389): s0 = the primitive time derivatives from the TVD slopes
template <int NV>
PPK_HD void sources(const HSt<NV> &q, const HSt<NV> &dqX, const HSt<NV> &dqY, const HSt<NV> &dqZ,
                    double gamma0, HSt<NV> &s0) {
  const double r = q[ID], p = q[IP], u = q[IU], v = q[IV];
  const double ir = recip(r);
  if constexpr (NV == 5) {
    const double w = q[IW];
    s0[ID] = -u * dqX[ID] - v * dqY[ID] - w * dqZ[ID] - (dqX[IU] + dqY[IV] + dqZ[IW]) * r;
    s0[IP] = -u * dqX[IP] - v * dqY[IP] - w * dqZ[IP] - (dqX[IU] + dqY[IV] + dqZ[IW]) * gamma0 * p;
    s0[IU] = -u * dqX[IU] - v * dqY[IU] - w * dqZ[IU] - divr(dqX[IP], r, ir);
    s0[IV] = -u * dqX[IV] - v * dqY[IV] - w * dqZ[IV] - divr(dqY[IP], r, ir);
    s0[IW] = -u * dqX[IW] - v * dqY[IW] - w * dqZ[IW] - divr(dqZ[IP], r, ir);
  } else {
    s0[ID] = -u * dqX[ID] - v * dqY[ID] - (dqX[IU] + dqY[IV]) * r;
    s0[IP] = -u * dqX[IP] - v * dqY[IP] - (dqX[IU] + dqY[IV]) * gamma0 * p;
    s0[IU] = -u * dqX[IU] - v * dqY[IU] - divr(dqX[IP], r, ir);
    s0[IV] = -u * dqX[IV] - v * dqY[IV] - divr(dqY[IP], r, ir);
  }
}

// the state of a cell on one of its faces: q -/+ dq/2 (min / max face of the slope's
// direction) + s0 * dtd / 2, density floored.  "q - HALF_F * d" of the reference is q + (-0.5 * d)
// here, the same IEEE result.
template <int NV>
PPK_HD void face_state(const HSt<NV> &q, const HSt<NV> &dq, const HSt<NV> &s0, double dtd,
                       bool max_side, double smallr, HSt<NV> &qface) {
  const double h = max_side ? 0.5 : -0.5;
#pragma unroll
  for (int n = 0; n < NV; ++n) qface[n] = q[n] + h * dq[n] + s0[n] * dtd * 0.5;
  qface[ID] = fmax(smallr, qface[ID]);
}

// cmpflx (RiemannSolvers.h:20-58)
template <int NV>
PPK_HD void cmpflx(const HSt<NV> &qg, HSt<NV> &flux, const HydroPar &P) {
  flux[ID] = qg[ID] * qg[IU];
  flux[IU] = flux[ID] * qg[IU] + qg[IP];
  flux[IV] = flux[ID] * qg[IV];
  if constexpr (NV == 5) flux[IW] = flux[ID] * qg[IW];
  const double entho = P.entho;
  double ekin = 0.5 * qg[ID] * (qg[IU] * qg[IU] + qg[IV] * qg[IV]);
  if constexpr (NV == 5) ekin += 0.5 * qg[ID] * (qg[IW] * qg[IW]);
  const double etot = qg[IP] * entho + ekin;
  flux[IP] = qg[IU] * (etot + qg[IP]);
}

// riemann_approx (RiemannSolvers.h:60-216): two-shock approximation, Newton iterations
template <int NV>
PPK_HD void riemann_approx(const HSt<NV> &ql, const HSt<NV> &qr, HSt<NV> &flux,
                           const HydroPar &P) {
  const double gamma0 = P.gamma0, gamma6 = P.gamma6, smallr = P.smallr, smallc = P.smallc;
  const double smallp = P.smallp, smallpp = P.smallpp;
  const double rl = fmax(ql[ID], smallr), ul = ql[IU], pl = fmax(ql[IP], rl * smallp);
  const double rr = fmax(qr[ID], smallr), ur = qr[IU], pr = fmax(qr[IP], rr * smallp);
  const double cl = gamma0 * pl * rl;
  const double cr = gamma0 * pr * rr;
  const double ipl = recip(pl), ipr = recip(pr);
  double wl = msqrt(cl);
  double wr = msqrt(cr);
  double pstar = fmax(divr((wr * pl + wl * pr) + wl * wr * (ul - ur), wl + wr, recip(wl + wr)), 0.0);
  double pold = pstar;
  double conv = 1.0;
  for (int iter = 0; (iter < 10) && (conv > 1e-6); ++iter) {
    const double wwl = msqrt(cl * (1.0 + divr(gamma6 * (pold - pl), pl, ipl)));
    const double wwr = msqrt(cr * (1.0 + divr(gamma6 * (pold - pr), pr, ipr)));
    const double qql = divr(2.0 * wwl * wwl * wwl, wwl * wwl + cl, recip(wwl * wwl + cl));
    const double qqr = divr(2.0 * wwr * wwr * wwr, wwr * wwr + cr, recip(wwr * wwr + cr));
    const double usl = ul - divr(pold - pl, wwl, recip(wwl));
    const double usr = ur + divr(pold - pr, wwr, recip(wwr));
    const double delp = fmax(divr(qqr * qql, qqr + qql, recip(qqr + qql)) * (usl - usr), -pold);
    pold = pold + delp;
    conv = fabs(divr(delp, pold + smallpp, recip(pold + smallpp)));
  }
  pstar = pold;
  wl = msqrt(cl * (1.0 + divr(gamma6 * (pstar - pl), pl, ipl)));
  wr = msqrt(cr * (1.0 + divr(gamma6 * (pstar - pr), pr, ipr)));
  const double ustar =
      0.5 * (ul + divr(pl - pstar, wl, recip(wl)) + ur - divr(pr - pstar, wr, recip(wr)));
  const double sgnm = copysign(1.0, ustar);
  double ro, uo, po, wo;
  if (sgnm > 0.0) {
    ro = rl; uo = ul; po = pl; wo = wl;
  } else {
    ro = rr; uo = ur; po = pr; wo = wr;
  }
  const double iro = recip(ro);
  const double co = fmax(smallc, msqrt(fabs(divr(gamma0 * po, ro, iro))));
  const double den = 1.0 + divr(ro * (po - pstar), wo * wo, recip(wo * wo));
  const double rstar = fmax(divr(ro, den, recip(den)), smallr);
  const double cstar = fmax(smallc, msqrt(fabs(divr(gamma0 * pstar, rstar, recip(rstar)))));
  double spout = co - sgnm * uo;
  double spin = cstar - sgnm * ustar;
  const double ushock = divr(wo, ro, iro) - sgnm * uo;
  if (pstar >= po) {
    spin = ushock;
    spout = ushock;
  }
  const double scr = fmax(spout - spin, smallc + fabs(spout + spin));
  double frac = 0.5 * (1.0 + divr(spout + spin, scr, recip(scr)));
  if (frac != frac)
    frac = 0.0;
  else
    frac = frac >= 1.0 ? 1.0 : frac <= 0.0 ? 0.0 : frac;
  HSt<NV> qg;
  qg[ID] = frac * rstar + (1.0 - frac) * ro;
  qg[IU] = frac * ustar + (1.0 - frac) * uo;
  qg[IP] = frac * pstar + (1.0 - frac) * po;
  if (spout < 0.0) {
    qg[ID] = ro; qg[IU] = uo; qg[IP] = po;
  }
  if (spin > 0.0) {
    qg[ID] = rstar; qg[IU] = ustar; qg[IP] = pstar;
  }
  if (sgnm > 0.0) {
    qg[IV] = ql[IV];
    if constexpr (NV == 5) qg[IW] = ql[IW];
  } else {
    qg[IV] = qr[IV];
    if constexpr (NV == 5) qg[IW] = qr[IW];
  }
  cmpflx(qg, flux, P);
}

// the conservative states and physical fluxes LLF and HLL build (RiemannSolvers.h:218-442)
template <int NV>
PPK_HD void cons_and_flux(const HSt<NV> &q, double entho, bool hll, HSt<NV> &u, HSt<NV> &f) {
  u[ID] = q[ID];
  u[IP] = q[IP] * entho + 0.5 * q[ID] * q[IU] * q[IU];
  u[IP] += 0.5 * q[ID] * q[IV] * q[IV];
  if constexpr (NV == 5) u[IP] += 0.5 * q[ID] * q[IW] * q[IW];
  u[IU] = q[ID] * q[IU];
  u[IV] = q[ID] * q[IV];
  if constexpr (NV == 5) u[IW] = q[ID] * q[IW];
  f[ID] = hll ? u[IU] : u[ID] * q[IU];
  f[IP] = q[IU] * (u[IP] + q[IP]);
  f[IU] = q[IP] + u[IU] * q[IU];
  f[IV] = f[ID] * q[IV];
  if constexpr (NV == 5) f[IW] = f[ID] * q[IW];
}

template <int NV>
PPK_HD void riemann_llf(const HSt<NV> &ql, const HSt<NV> &qr, HSt<NV> &flux, const HydroPar &P) {
  const double gamma0 = P.gamma0, smallr = P.smallr, smallp = P.smallp;
  const double entho = P.entho;
  const double rl = fmax(ql[ID], smallr), ul = ql[IU], pl = fmax(ql[IP], rl * smallp);
  const double rr = fmax(qr[ID], smallr), ur = qr[IU], pr = fmax(qr[IP], rr * smallp);
  const double cl = msqrt(divr(gamma0 * pl, rl, recip(rl)));
  const double cr = msqrt(divr(gamma0 * pr, rr, recip(rr)));
  const double cmax = fmax(fabs(ul) + cl, fabs(ur) + cr);
  HSt<NV> uleft, uright, fleft, fright;
  cons_and_flux(ql, entho, false, uleft, fleft);
  cons_and_flux(qr, entho, false, uright, fright);
#pragma unroll
  for (int n = 0; n < NV; ++n)
    flux[n] = 0.5 * (fleft[n] + fright[n] - cmax * (uright[n] - uleft[n]));
}

template <int NV>
PPK_HD void riemann_hll(const HSt<NV> &ql, const HSt<NV> &qr, HSt<NV> &flux, const HydroPar &P) {
  const double gamma0 = P.gamma0, smallr = P.smallr, smallp = P.smallp;
  const double entho = P.entho;
  const double rl = fmax(ql[ID], smallr), ul = ql[IU], pl = fmax(ql[IP], rl * smallp);
  const double rr = fmax(qr[ID], smallr), ur = qr[IU], pr = fmax(qr[IP], rr * smallp);
  const double cl = msqrt(divr(gamma0 * pl, rl, recip(rl)));
  const double cr = msqrt(divr(gamma0 * pr, rr, recip(rr)));
  const double SL = fmin(fmin(ul, ur) - fmax(cl, cr), 0.0);
  const double SR = fmax(fmax(ul, ur) + fmax(cl, cr), 0.0);
  HSt<NV> uleft, uright, fleft, fright;
  cons_and_flux(ql, entho, true, uleft, fleft);
  cons_and_flux(qr, entho, true, uright, fright);
  const double iw = recip(SR - SL);
#pragma unroll
  for (int n = 0; n < NV; ++n)
    flux[n] = divr(SR * fleft[n] - SL * fright[n] + SR * SL * (uright[n] - uleft[n]), SR - SL, iw);
}

template <int NV>
PPK_HD void riemann_hllc(const HSt<NV> &ql, const HSt<NV> &qr, HSt<NV> &flux, const HydroPar &P) {
  const double gamma0 = P.gamma0, smallr = P.smallr, smallp = P.smallp, smallc = P.smallc;
  const double entho = P.entho;
  const double rl = fmax(ql[ID], smallr);
  const double pl = fmax(ql[IP], rl * smallp);
  const double ul = ql[IU];
  double ecinl = 0.5 * rl * ul * ul;
  ecinl += 0.5 * rl * ql[IV] * ql[IV];
  if constexpr (NV == 5) ecinl += 0.5 * rl * ql[IW] * ql[IW];
  const double etotl = pl * entho + ecinl;
  const double ptotl = pl;
  const double rr = fmax(qr[ID], smallr);
  const double pr = fmax(qr[IP], rr * smallp);
  const double ur = qr[IU];
  double ecinr = 0.5 * rr * ur * ur;
  ecinr += 0.5 * rr * qr[IV] * qr[IV];
  if constexpr (NV == 5) ecinr += 0.5 * rr * qr[IW] * qr[IW];
  const double etotr = pr * entho + ecinr;
  const double ptotr = pr;
  const double cfastl = msqrt(fmax(divr(gamma0 * pl, rl, recip(rl)), smallc * smallc));
  const double cfastr = msqrt(fmax(divr(gamma0 * pr, rr, recip(rr)), smallc * smallc));
  const double SL = fmin(ul, ur) - fmax(cfastl, cfastr);
  const double SR = fmax(ul, ur) + fmax(cfastl, cfastr);
  const double rcl = rl * (ul - SL);
  const double rcr = rr * (SR - ur);
  const double isum = recip(rcr + rcl);
  const double ustar = divr(rcr * ur + rcl * ul + (ptotl - ptotr), rcr + rcl, isum);
  const double ptotstar = divr(rcr * ptotl + rcl * ptotr + rcl * rcr * (ul - ur), rcr + rcl, isum);
  const double isl = recip(SL - ustar), isr = recip(SR - ustar);
  const double rstarl = divr(rl * (SL - ul), SL - ustar, isl);
  const double etotstarl = divr((SL - ul) * etotl - ptotl * ul + ptotstar * ustar, SL - ustar, isl);
  const double rstarr = divr(rr * (SR - ur), SR - ustar, isr);
  const double etotstarr = divr((SR - ur) * etotr - ptotr * ur + ptotstar * ustar, SR - ustar, isr);
  double ro, uo, ptoto, etoto;
  if (SL > 0.0) {
    ro = rl; uo = ul; ptoto = ptotl; etoto = etotl;
  } else if (ustar > 0.0) {
    ro = rstarl; uo = ustar; ptoto = ptotstar; etoto = etotstarl;
  } else if (SR > 0.0) {
    ro = rstarr; uo = ustar; ptoto = ptotstar; etoto = etotstarr;
  } else {
    ro = rr; uo = ur; ptoto = ptotr; etoto = etotr;
  }
  flux[ID] = ro * uo;
  flux[IU] = ro * uo * uo + ptoto;
  flux[IP] = (etoto + ptoto) * uo;
  if (flux[ID] > 0.0)
    flux[IV] = flux[ID] * ql[IV];
  else
    flux[IV] = flux[ID] * qr[IV];
  if constexpr (NV == 5) {
    if (flux[ID] > 0.0)
      flux[IW] = flux[ID] * ql[IW];
    else
      flux[IW] = flux[ID] * qr[IW];
  }
}

// riemann_hydro (RiemannSolvers.h:573-633) with the solver fixed at compile time (RS: enum
// ppk_riemann), so each kernel holds one solver and its flux stays in registers
template <int RS, int NV>
PPK_HD void riemann_rs(const HSt<NV> &ql, const HSt<NV> &qr, HSt<NV> &flux, const HydroPar &P) {
  if constexpr (RS == PPK_RIEMANN_APPROX)
    riemann_approx(ql, qr, flux, P);
  else if constexpr (RS == PPK_RIEMANN_HLL)
    riemann_hll(ql, qr, flux, P);
  else if constexpr (RS == PPK_RIEMANN_HLLC)
    riemann_hllc(ql, qr, flux, P);
  else
    riemann_llf(ql, qr, flux, P);
}

// riemann_hydro with the solver chosen at run time; HLL exists in 2D only (3D rejects it at
// create)
template <int NV>
PPK_HD void riemann(const HSt<NV> &ql, const HSt<NV> &qr, HSt<NV> &flux, const HydroPar &P) {
  if (P.riemann == PPK_RIEMANN_APPROX)
    riemann_approx(ql, qr, flux, P);
  else if (P.riemann == PPK_RIEMANN_HLL)
    riemann_hll(ql, qr, flux, P);
  else if (P.riemann == PPK_RIEMANN_HLLC)
    riemann_hllc(ql, qr, flux, P);
  else
    riemann_llf(ql, qr, flux, P);
}

}  // namespace hyd
}  // namespace ppk
