// Device-side MHD building blocks of the MUSCL-Hancock + HLLD + CT step (gfx950).
//
// Every function is an inline on values held in registers (no local arrays indexed
// at run time, so nothing spills to scratch).  Arithmetic is written in the operand order of
// the reference so that a build with -ffp-contract=off reproduces the reference's double
// results bit for bit; the default build lets the compiler contract a*b+c into FMA.
//
// Variable order ID, IP, IU, IV, IW, IA, IB, IC: reference src/shared/enums.h:17-34.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "ppkmhd_amd.h"

// host+device inline: the kernels' per-thread code also builds for the CPU (index checks)
#define PPK_HD __host__ __device__ __forceinline__

namespace ppk {

enum { ID = 0, IP = 1, IU = 2, IV = 3, IW = 4, IA = 5, IB = 6, IC = 7 };
enum { IX = 0, IY = 1, IZ = 2 };
enum { IRT = 0, IRB = 1, ILT = 2, ILB = 3 };
enum { EMFX = 0, EMFY = 1, EMFZ = 2 };
enum { I_EMFZ = 0, I_EMFY = 1, I_EMFX = 2 };

// POD copy of the solver parameters passed by value to every kernel (as the reference passes
// HydroParams by value into every functor).
struct DevParams {
  int dim, nx, ny, nz, gw;
  int isize, jsize, ksize;
  int riemann;
  int bc[6];
  int slope12;          // slope_type is 1 or 2 (hydro slopes active)
  long long plane;      // isize * jsize
  long long vstride;    // isize * jsize * ksize (distance between variables)
  double dx, dy, dz;
  double gamma0, slope_type, smallr, smallc, smallp, cIso, cfl;
  int tile3;            // 3D half-step tile shape (batch tables; 0 for single grids)
  int brick;            // block-to-cells shape of the fused per-cell kernels (batch tables)
  int bulk32;           // FP32-bulk form: half step, records, face/edge solves in float
  int march;            // 3D single-pass step: blocks of its launch (0: not used)
};

template <class R>
struct St {
  R v[8];
  PPK_HD R &operator[](int i) { return v[i]; }
  PPK_HD const R &operator[](int i) const { return v[i]; }
};

template <class R>
PPK_HD R dmax(R a, R b) { return fmax(a, b); }
// |x| in the argument's own type (host fabs(float) would widen to double)
PPK_HD double dabs(double x) { return __builtin_fabs(x); }
PPK_HD float dabs(float x) { return __builtin_fabsf(x); }
// sign bit test (true for -0 and negative values), host and device
template <class R>
PPK_HD bool sign_bit(R x) { return __builtin_signbit(x); }

// Division and square root.  Built with PPK_EXACT_DIV (the strict variant) every a / b and
// sqrt(x) is the IEEE operation as written.  Otherwise (the default FP64 variant):
//  - a denominator used several times is inverted once and each quotient is a * (1/b);
//  - on the GPU, 1/b is v_rcp_f64 refined by one third-order step, sqrt(x) is v_rsq_f64
//    refined and corrected with its residual, and a / sqrt(r) is a * (refined rsq).  On
//    MI355X, over log-uniform inputs in [2^-40, 2^40] (ppk_mhd_math_selfcheck), the refined
//    reciprocal and square root equal the IEEE results and the inverse square root is within
//    1 ulp; each costs 4-8 instructions where the IEEE sequences with range scaling take
//    11-14.  Inputs here are positive normal numbers (densities, pressures,
//    wave-speed differences); sqrt(0) is 0, other zero and non-finite inputs are not
//    handled like IEEE.
#ifdef PPK_EXACT_DIV
constexpr bool kSharedRecip = false;
#else
constexpr bool kSharedRecip = true;
#endif
#if defined(__HIP_DEVICE_COMPILE__) && !defined(PPK_EXACT_DIV)
// one third-order step from the hardware estimate r0 (relative error eps): with
// e = 1 - b r0, 1/b = r0 (1 + e + e^2 + ...); with e = 1 - x y0^2, 1/sqrt(x) =
// y0 (1 + e/2 + 3e^2/8 + ...): the truncation error is O(eps^3), below the final rounding
PPK_HD double fast_rcp(double b) {
  const double r = __builtin_amdgcn_rcp(b);
  const double e = fma(-b, r, 1.0);
  return fma(r, fma(e, e, e), r);
}
PPK_HD double fast_rsq(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  const double e = fma(-x * y, y, 1.0);
  return fma(y * e, fma(e, 0.375, 0.5), y);
}
PPK_HD double fast_sqrt(double x) {
  // Goldschmidt: g ~ sqrt(x) and h ~ 1/(2 sqrt(x)) refined together (r = 1/2 - g h), then
  // one correction with the residual x - s^2.  sqrt(+-0) = +-0 (the magnetosonic term
  // d2^2 - c2 bn^2 / d can vanish): rsq(0) = inf is clamped to the largest double, so that
  // g = x q = 0 instead of NaN (and every step below keeps the zero); positive inputs are
  // unaffected, a NaN stays NaN
  const double q = fmin(__builtin_amdgcn_rsq(fabs(x)), 1.7976931348623157e308);
  const double h = 0.5 * q;
  const double g = x * q;
  const double r = fma(-g, h, 0.5);
  const double s = fma(g, r, g);
  const double d = fma(-s, s, x);
  return fma(d, fma(h, r, h), s);
}
#else
PPK_HD double fast_rcp(double b) { return 1.0 / b; }
PPK_HD double fast_rsq(double x) { return 1.0 / sqrt(x); }
PPK_HD double fast_sqrt(double x) { return sqrt(x); }
#endif
template <class R>
PPK_HD R recip(R b) { return kSharedRecip ? R(fast_rcp(double(b))) : R(0); }
template <class R>
PPK_HD R divr(R a, R b, R inv_b) { return kSharedRecip ? a * inv_b : a / b; }
template <class R>
PPK_HD R msqrt(R x) { return kSharedRecip ? R(fast_sqrt(double(x))) : sqrt(x); }
// a / sqrt(r)
template <class R>
PPK_HD R over_sqrt(R a, R r) { return kSharedRecip ? a * R(fast_rsq(double(r))) : a / sqrt(r); }
// float (FP32-bulk form): single-instruction reciprocal, sqrt and rsq on the GPU (about
// 1 ulp of float, far inside that form's error budget); IEEE float elsewhere
#if defined(__HIP_DEVICE_COMPILE__) && !defined(PPK_EXACT_DIV)
PPK_HD float recip(float b) { return __builtin_amdgcn_rcpf(b); }
PPK_HD float msqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
PPK_HD float over_sqrt(float a, float r) { return a * __builtin_amdgcn_rsqf(r); }
#else
PPK_HD float recip(float b) { return 1.0f / b; }
PPK_HD float msqrt(float x) { return sqrtf(x); }
PPK_HD float over_sqrt(float a, float r) { return a / sqrtf(r); }
#endif
template <class R>
PPK_HD R dmin(R a, R b) { return fmin(a, b); }
// The magnetosonic discriminant d2^2 - c2 bn^2 / rho is >= 0 in exact arithmetic and 0 when
// the field is along the normal; rounded (and, in the default builds, FMA-contracted or float)
// it can come out a few ulp below 0, and its square root NaN.  The default builds clamp it at
// 0 (with fmax, one instruction, which like a compare and select gives 0 for a NaN input);
// the strict build keeps the reference's expression as written.
template <class R>
PPK_HD R nonneg(R x) { return kSharedRecip ? dmax(x, R(0)) : x; }

// compare-and-select maxima used by the 2D magnetic Riemann solver (FMAX4/FMIN4/FMAX5,
// reference src/shared/mhd_utils.h:36-77).  The default builds take fmax / fmin, one
// instruction each instead of a compare and two selects; for operands that are not NaN the
// value is the same: the maxima of fast and Alfven speeds have operands >= +0 only, and the
// extrema of velocities are offset by a positive fast speed (S = min u - max c_f, max u +
// max c_f), so the sign of a zero extremum does not reach the result.
template <class R>
PPK_HD R max4(R a0, R a1, R a2, R a3) {
  if (kSharedRecip) return dmax(dmax(dmax(a0, a1), a2), a3);
  R r = a0;
  r = (a1 > r) ? a1 : r;
  r = (a2 > r) ? a2 : r;
  r = (a3 > r) ? a3 : r;
  return r;
}
template <class R>
PPK_HD R min4(R a0, R a1, R a2, R a3) {
  if (kSharedRecip) return dmin(dmin(dmin(a0, a1), a2), a3);
  R r = a0;
  r = (a1 < r) ? a1 : r;
  r = (a2 < r) ? a2 : r;
  r = (a3 < r) ? a3 : r;
  return r;
}
template <class R>
PPK_HD R max5(R a0, R a1, R a2, R a3, R a4) {
  R r = max4(a0, a1, a2, a3);
  if (kSharedRecip) return dmax(r, a4);
  r = (a4 > r) ? a4 : r;
  return r;
}

// fast magnetosonic speed along dir, reference src/shared/mhd_utils.h:89-117
template <class R>
PPK_HD R speed_fast(const St<R> &q, int dir, R gamma0) {
  R d = q[ID], p = q[IP], a = q[IA], b = q[IB], c = q[IC];
  const R id = recip(d);
  R b2 = a * a + b * b + c * c;
  R c2 = divr(gamma0 * p, d, id);
  R d2 = R(0.5) * (divr(b2, d, id) + c2);
  R bn = dir == IX ? a : (dir == IY ? b : c);
  // c2 bn^2 / d: in the default builds (c2 / d) bn^2, c2 / d shared by the directions of a state
  const R cbd = kSharedRecip ? (c2 * id) * (bn * bn) : c2 * bn * bn / d;
  return msqrt(d2 + msqrt(nonneg(d2 * d2 - cbd)));
}

// per-cell CFL term sum_d (c_f,d + |v_d|)/dx_d, reference src/shared/mhd_utils.h:319-366 and
// src/muscl/MHDRunFunctors{2D,3D}.h:73/76
// In the default build the quotients are products with shared reciprocals; wherever both
// operands of a sum would then be products, the fused multiply-add is written out, so that
// every kernel that inlines this function rounds it the same way (the CFL dt of a step must
// not depend on which kernel computed it: the folded maximum of the update and the
// standalone pass agree bit for bit).
PPK_HD double inv_dt_cell(const St<double> &q, const DevParams &p) {
  const double gamma0 = p.gamma0;
  double d = q[ID], pp = q[IP], a = q[IA], b = q[IB], c = q[IC];
  double b2 = a * a + b * b + c * c;
  const double id = recip(d);
  double c2 = divr(gamma0 * pp, d, id);
  double d2 = 0.5 * (kSharedRecip ? fma(b2, id, c2) : b2 / d + c2);
  auto disc = [&](double bn) {  // d2^2 - c2 bn^2 / d
    return kSharedRecip ? fma(d2, d2, -(c2 * bn * bn * id)) : d2 * d2 - c2 * bn * bn / d;
  };
  double cf = msqrt(d2 + msqrt(nonneg(disc(a))));
  double vx = cf + dabs(q[IU]);
  cf = msqrt(d2 + msqrt(nonneg(disc(b))));
  double vy = cf + dabs(q[IV]);
  if (p.dim == 3) {
    cf = msqrt(d2 + msqrt(nonneg(disc(c))));
    double vz = cf + dabs(q[IW]);
    if (kSharedRecip)
      return fma(vz, recip(p.dz), fma(vy, recip(p.dy), vx * recip(p.dx)));
    return vx / p.dx + vy / p.dy + vz / p.dz;
  }
  if (kSharedRecip) return fma(vy, recip(p.dy), vx * recip(p.dx));
  return vx / p.dx + vy / p.dy;
}

// conservative -> primitive with the face-centred field averaged to the cell centre,
// reference src/muscl/MHDBaseFunctor3D.h:192-242
template <class R>
PPK_HD void constoprim(const St<R> &u, R magx, R magy, R magz, St<R> &q,
                                           const DevParams &p) {
  q[ID] = dmax(u[ID], R(p.smallr));
  const R ir = recip(q[ID]);
  q[IU] = divr(u[IU], q[ID], ir);
  q[IV] = divr(u[IV], q[ID], ir);
  q[IW] = divr(u[IW], q[ID], ir);
  q[IA] = R(0.5) * (u[IA] + magx);
  q[IB] = R(0.5) * (u[IB] + magy);
  q[IC] = R(0.5) * (u[IC] + magz);
  R eken = R(0.5) * (q[IU] * q[IU] + q[IV] * q[IV] + q[IW] * q[IW]);
  R emag = R(0.5) * (q[IA] * q[IA] + q[IB] * q[IB] + q[IC] * q[IC]);
  if (p.cIso > 0) {
    q[IP] = q[ID] * R(p.cIso) * R(p.cIso);
  } else {
    R eint = divr(u[IP] - emag, q[ID], ir) - eken;
    q[IP] = dmax((R(p.gamma0) - R(1.0)) * q[ID] * eint, q[ID] * R(p.smallp));
  }
}

// minmod (slope_type 1) / MC (slope_type 2) limited slope,
// reference src/muscl/MHDBaseFunctor2D.h:308-327.  The default builds give the magnitude
// the sign of dcen with copysign (one instruction for the compare, select and multiply by
// +-1); the value is the same, only a zero slope at dcen = -0 keeps the sign of dcen.
template <class R>
PPK_HD R limited_slope(R st, R q, R qplus, R qminus) {
  R dlft = st * (q - qminus);
  R drgt = st * (qplus - q);
  R dcen = R(0.5) * (qplus - qminus);
  R slop = dmin(dabs(dlft), dabs(drgt));
  R dlim = slop;
  if ((dlft * drgt) <= R(0)) dlim = R(0);
  if (kSharedRecip) return copysign(dmin(dlim, dabs(dcen)), dcen);
  R dsgn = (dcen >= R(0)) ? R(1) : R(-1);
  return dsgn * dmin(dlim, dabs(dcen));
}

// ------------------------------------------------------------------------------------------
// 1D Riemann solvers, reference src/shared/RiemannSolvers_MHD.h.  Inputs are rotated so that
// IU / IA are normal to the face; the flux keeps that rotation.
// ------------------------------------------------------------------------------------------
// HLLD, Miyoshi & Kusano (2005) as written in the reference (:133-367)
template <class R>
PPK_HD void riemann_hlld(St<R> &qleft, St<R> &qright, St<R> &flux,
                                             const DevParams &par) {
  const R gamma0 = R(par.gamma0);
  const R entho = kSharedRecip ? recip(gamma0 - R(1.0)) : R(1.0) / (gamma0 - R(1.0));

  R a = R(0.5) * (qleft[IA] + qright[IA]);
  R sgnm = (a >= R(0)) ? R(1) : R(-1);
  qleft[IA] = a;
  qright[IA] = a;
  if (par.cIso > 0) {
    qleft[IP] = qleft[ID] * R(par.cIso) * R(par.cIso);
    qright[IP] = qright[ID] * R(par.cIso) * R(par.cIso);
  }

  R rl = qleft[ID], pl = qleft[IP], ul = qleft[IU], vl = qleft[IV], wl = qleft[IW];
  R bl = qleft[IB], cl = qleft[IC];
  R ecinl = R(0.5) * (ul * ul + vl * vl + wl * wl) * rl;
  R emagl = R(0.5) * (a * a + bl * bl + cl * cl);
  R etotl = pl * entho + ecinl + emagl;
  R ptotl = pl + emagl;
  R vdotbl = ul * a + vl * bl + wl * cl;

  R rr = qright[ID], pr = qright[IP], ur = qright[IU], vr = qright[IV], wr = qright[IW];
  R br = qright[IB], cr = qright[IC];
  R ecinr = R(0.5) * (ur * ur + vr * vr + wr * wr) * rr;
  R emagr = R(0.5) * (a * a + br * br + cr * cr);
  R etotr = pr * entho + ecinr + emagr;
  R ptotr = pr + emagr;
  R vdotbr = ur * a + vr * br + wr * cr;

  R cfastl = speed_fast(qleft, IX, gamma0);
  R cfastr = speed_fast(qright, IX, gamma0);

  R sl = dmin(ul, ur) - dmax(cfastl, cfastr);
  R sr = dmax(ul, ur) + dmax(cfastl, cfastr);

  R rcl = rl * (ul - sl);
  R rcr = rr * (sr - ur);

  const R irc = recip(rcr + rcl);
  R ustar = divr(rcr * ur + rcl * ul + (ptotl - ptotr), rcr + rcl, irc);
  R ptotstar = divr(rcr * ptotl + rcl * ptotr + rcl * rcr * (ul - ur), rcr + rcl, irc);

  // the degenerate-state test |estar / a^2 - 1| <= 1e-8, in the default builds as
  // |estar - a^2| <= 1e-8 a^2 (no reciprocal; the same test but for rounding at its edge)
  auto degenerate = [&](R estar) {
    return a * a > R(0) && (kSharedRecip ? dabs(estar - a * a) <= R(1e-8) * (a * a)
                                         : dabs(estar / (a * a) - R(1)) <= R(1e-8));
  };
  const R isl = recip(sl - ustar);
  R rstarl = divr(rl * (sl - ul), sl - ustar, isl);
  R estar = rl * (sl - ul) * (sl - ustar) - a * a;
  R el = rl * (sl - ul) * (sl - ul) - a * a;
  R vstarl, wstarl, bstarl, cstarl;
  if (degenerate(estar)) {
    vstarl = vl;
    bstarl = bl;
    wstarl = wl;
    cstarl = cl;
  } else {
    const R ie = recip(estar);
    vstarl = vl - divr(a * bl * (ustar - ul), estar, ie);
    bstarl = divr(bl * el, estar, ie);
    wstarl = wl - divr(a * cl * (ustar - ul), estar, ie);
    cstarl = divr(cl * el, estar, ie);
  }
  R vdotbstarl = ustar * a + vstarl * bstarl + wstarl * cstarl;
  R etotstarl = divr((sl - ul) * etotl - ptotl * ul + ptotstar * ustar + a * (vdotbl - vdotbstarl),
                     sl - ustar, isl);
  R sqrrstarl = msqrt(rstarl);
  R calfvenl = over_sqrt(dabs(a), rstarl);
  R sal = ustar - calfvenl;

  const R isr = recip(sr - ustar);
  R rstarr = divr(rr * (sr - ur), sr - ustar, isr);
  estar = rr * (sr - ur) * (sr - ustar) - a * a;
  R er = rr * (sr - ur) * (sr - ur) - a * a;
  R vstarr, wstarr, bstarr, cstarr;
  if (degenerate(estar)) {
    vstarr = vr;
    bstarr = br;
    wstarr = wr;
    cstarr = cr;
  } else {
    const R ie = recip(estar);
    vstarr = vr - divr(a * br * (ustar - ur), estar, ie);
    bstarr = divr(br * er, estar, ie);
    wstarr = wr - divr(a * cr * (ustar - ur), estar, ie);
    cstarr = divr(cr * er, estar, ie);
  }
  R vdotbstarr = ustar * a + vstarr * bstarr + wstarr * cstarr;
  R etotstarr = divr((sr - ur) * etotr - ptotr * ur + ptotstar * ustar + a * (vdotbr - vdotbstarr),
                     sr - ustar, isr);
  R sqrrstarr = msqrt(rstarr);
  R calfvenr = over_sqrt(dabs(a), rstarr);
  R sar = ustar + calfvenr;

  const R isq = recip(sqrrstarl + sqrrstarr);
  R vstarstar = divr(sqrrstarl * vstarl + sqrrstarr * vstarr + sgnm * (bstarr - bstarl),
                     sqrrstarl + sqrrstarr, isq);
  R wstarstar = divr(sqrrstarl * wstarl + sqrrstarr * wstarr + sgnm * (cstarr - cstarl),
                     sqrrstarl + sqrrstarr, isq);
  R bstarstar = divr(sqrrstarl * bstarr + sqrrstarr * bstarl +
                         sgnm * sqrrstarl * sqrrstarr * (vstarr - vstarl),
                     sqrrstarl + sqrrstarr, isq);
  R cstarstar = divr(sqrrstarl * cstarr + sqrrstarr * cstarl +
                         sgnm * sqrrstarl * sqrrstarr * (wstarr - wstarl),
                     sqrrstarl + sqrrstarr, isq);
  R vdotbstarstar = ustar * a + vstarstar * bstarstar + wstarstar * cstarstar;
  R etotstarstarl = etotstarl - sgnm * sqrrstarl * (vdotbstarl - vdotbstarstar);
  R etotstarstarr = etotstarr + sgnm * sqrrstarr * (vdotbstarr - vdotbstarstar);

  R ro, uo, vo, wo, bo, co, ptoto, etoto, vdotbo;
  if (sl > R(0)) {
    ro = rl; uo = ul; vo = vl; wo = wl; bo = bl; co = cl;
    ptoto = ptotl; etoto = etotl; vdotbo = vdotbl;
  } else if (sal > R(0)) {
    ro = rstarl; uo = ustar; vo = vstarl; wo = wstarl; bo = bstarl; co = cstarl;
    ptoto = ptotstar; etoto = etotstarl; vdotbo = vdotbstarl;
  } else if (ustar > R(0)) {
    ro = rstarl; uo = ustar; vo = vstarstar; wo = wstarstar; bo = bstarstar; co = cstarstar;
    ptoto = ptotstar; etoto = etotstarstarl; vdotbo = vdotbstarstar;
  } else if (sar > R(0)) {
    ro = rstarr; uo = ustar; vo = vstarstar; wo = wstarstar; bo = bstarstar; co = cstarstar;
    ptoto = ptotstar; etoto = etotstarstarr; vdotbo = vdotbstarstar;
  } else if (sr > R(0)) {
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
  flux[IA] = R(0);
  flux[IB] = bo * uo - a * vo;
  flux[IC] = co * uo - a * wo;
}

// ideal MHD flux and conservative state of a primitive state, reference
// src/shared/mhd_utils.h:173-231 (used by HLL and LLF)
template <class R>
PPK_HD void mhd_flux(const St<R> &q, St<R> &cvar, St<R> &ff,
                                         const DevParams &par) {
  R gamma0 = R(par.gamma0);
  R p = par.cIso > 0 ? q[ID] * R(par.cIso) * R(par.cIso) : q[IP];
  R entho = kSharedRecip ? recip(gamma0 - R(1.0)) : R(1.0) / (gamma0 - R(1.0));
  R d = q[ID], u = q[IU], v = q[IV], w = q[IW], a = q[IA], b = q[IB], c = q[IC];
  R ecin = R(0.5) * (u * u + v * v + w * w) * d;
  R emag = R(0.5) * (a * a + b * b + c * c);
  R etot = p * entho + ecin + emag;
  R ptot = p + emag;
  cvar[ID] = d; cvar[IP] = etot; cvar[IU] = d * u; cvar[IV] = d * v; cvar[IW] = d * w;
  cvar[IA] = a; cvar[IB] = b; cvar[IC] = c;
  ff[ID] = d * u;
  ff[IP] = (etot + ptot) * u - a * (a * u + b * v + c * w);
  ff[IU] = d * u * u - a * a + ptot;
  ff[IV] = d * u * v - a * b;
  ff[IW] = d * u * w - a * c;
  ff[IA] = R(0);
  ff[IB] = b * u - a * v;
  ff[IC] = c * u - a * w;
}

// HLL, reference src/shared/RiemannSolvers_MHD.h:27-67
template <class R>
PPK_HD void riemann_hll(St<R> &ql, St<R> &qr, St<R> &flux,
                                            const DevParams &par) {
  R bx_mean = R(0.5) * (ql[IA] + qr[IA]);
  ql[IA] = bx_mean;
  qr[IA] = bx_mean;
  St<R> ul, fl, ur, fr;
  mhd_flux(ql, ul, fl, par);
  mhd_flux(qr, ur, fr, par);
  R cfl = speed_fast(ql, IX, R(par.gamma0));
  R cfr = speed_fast(qr, IX, R(par.gamma0));
  R sl = dmin(dmin(ql[IU], qr[IU]) - dmax(cfl, cfr), R(0));
  R sr = dmax(dmax(ql[IU], qr[IU]) + dmax(cfl, cfr), R(0));
#pragma unroll
  for (int v = 0; v < 8; ++v)
    flux[v] = (sr * fl[v] - sl * fr[v] + sr * sl * (ur[v] - ul[v])) / (sr - sl);
}

// LLF, reference src/shared/RiemannSolvers_MHD.h:81-111
template <class R>
PPK_HD void riemann_llf(St<R> &ql, St<R> &qr, St<R> &flux,
                                            const DevParams &par) {
  R bx_mean = R(0.5) * (ql[IA] + qr[IA]);
  ql[IA] = bx_mean;
  qr[IA] = bx_mean;
  St<R> ul, fl, ur, fr;
  mhd_flux(ql, ul, fl, par);
  mhd_flux(qr, ur, fr, par);
#pragma unroll
  for (int v = 0; v < 8; ++v) flux[v] = (fl[v] + fr[v]) / R(2);
  auto info = [&](const St<R> &q) {
    R d = q[ID], pp = q[IP], a = q[IA], b = q[IB], c = q[IC];
    R b2 = a * a + b * b + c * c;
    R c2 = R(par.gamma0) * pp / d;
    R d2 = R(0.5) * (b2 / d + c2);
    R cf = sqrt(d2 + sqrt(nonneg(d2 * d2 - c2 * a * a / d)));
    return cf + dabs(q[IU]);
  };
  R vel_info = dmax(info(ql), info(qr));
#pragma unroll
  for (int v = 0; v < 8; ++v) flux[v] = flux[v] - vel_info * (ur[v] - ul[v]) / R(2);
}

template <class R>
PPK_HD void riemann_mhd(St<R> &ql, St<R> &qr, St<R> &flux,
                                            const DevParams &p) {
  if (p.riemann == PPK_RIEMANN_HLLD)
    riemann_hlld(ql, qr, flux, p);
  else if (p.riemann == PPK_RIEMANN_HLL)
    riemann_hll(ql, qr, flux, p);
  else
    riemann_llf(ql, qr, flux, p);
}
// the same with the solver fixed at compile time (RS: enum ppk_riemann; RS < 0: p.riemann at
// run time), so a kernel can hold only the solver it runs
template <int RS, class R>
PPK_HD void riemann_mhd_rs(St<R> &ql, St<R> &qr, St<R> &flux, const DevParams &p) {
  if constexpr (RS == PPK_RIEMANN_HLLD)
    riemann_hlld(ql, qr, flux, p);
  else if constexpr (RS == PPK_RIEMANN_HLL)
    riemann_hll(ql, qr, flux, p);
  else if constexpr (RS == PPK_RIEMANN_LLF)
    riemann_llf(ql, qr, flux, p);
  else
    riemann_mhd(ql, qr, flux, p);
}

// ------------------------------------------------------------------------------------------
// 2D magnetic Riemann solver (EMF at a cell edge), reference src/shared/RiemannSolvers_MHD.h
// ------------------------------------------------------------------------------------------
// mag_riemann2d_hlld (:398-630), with the reference's branch-free 0/1 blend (:569-596): the
// region masks come from the sign bit (a -0 speed selects the "negative" side).
template <class R>
PPK_HD R mag_riemann2d_hlld(const St<R> &qLL, const St<R> &qRL,
                                               const St<R> &qLR, const St<R> &qRR, R ELL, R ERL,
                                               R ELR, R ERR, const DevParams &p) {
  const R gamma0 = R(p.gamma0);
  R rLL = qLL[ID], pLL = qLL[IP], uLL = qLL[IU], vLL = qLL[IV], aLL = qLL[IA], bLL = qLL[IB], cLL = qLL[IC];
  R rLR = qLR[ID], pLR = qLR[IP], uLR = qLR[IU], vLR = qLR[IV], aLR = qLR[IA], bLR = qLR[IB], cLR = qLR[IC];
  R rRL = qRL[ID], pRL = qRL[IP], uRL = qRL[IU], vRL = qRL[IV], aRL = qRL[IA], bRL = qRL[IB], cRL = qRL[IC];
  R rRR = qRR[ID], pRR = qRR[IP], uRR = qRR[IU], vRR = qRR[IV], aRR = qRR[IA], bRR = qRR[IB], cRR = qRR[IC];

  R cFastLLx = speed_fast(qLL, IX, gamma0);
  R cFastLRx = speed_fast(qLR, IX, gamma0);
  R cFastRLx = speed_fast(qRL, IX, gamma0);
  R cFastRRx = speed_fast(qRR, IX, gamma0);
  R cFastLLy = speed_fast(qLL, IY, gamma0);
  R cFastLRy = speed_fast(qLR, IY, gamma0);
  R cFastRLy = speed_fast(qRL, IY, gamma0);
  R cFastRRy = speed_fast(qRR, IY, gamma0);

  R SL = min4(uLL, uLR, uRL, uRR) - max4(cFastLLx, cFastLRx, cFastRLx, cFastRRx);
  R SR = max4(uLL, uLR, uRL, uRR) + max4(cFastLLx, cFastLRx, cFastRLx, cFastRRx);
  R SB = min4(vLL, vLR, vRL, vRR) - max4(cFastLLy, cFastLRy, cFastRLy, cFastRRy);
  R ST = max4(vLL, vLR, vRL, vRR) + max4(cFastLLy, cFastLRy, cFastRLy, cFastRRy);

  R PtotLL = pLL + R(0.5) * (aLL * aLL + bLL * bLL + cLL * cLL);
  R PtotLR = pLR + R(0.5) * (aLR * aLR + bLR * bLR + cLR * cLR);
  R PtotRL = pRL + R(0.5) * (aRL * aRL + bRL * bRL + cRL * cRL);
  R PtotRR = pRR + R(0.5) * (aRR * aRR + bRR * bRR + cRR * cRR);

  R rcLLx = rLL * (uLL - SL);
  R rcRLx = rRL * (SR - uRL);
  R rcLRx = rLR * (uLR - SL);
  R rcRRx = rRR * (SR - uRR);
  R rcLLy = rLL * (vLL - SB);
  R rcLRy = rLR * (ST - vLR);
  R rcRLy = rRL * (vRL - SB);
  R rcRRy = rRR * (ST - vRR);

  const R rcx = rcLLx + rcLRx + rcRLx + rcRRx, rcy = rcLLy + rcLRy + rcRLy + rcRRy;
  R ustar = divr(rcLLx * uLL + rcLRx * uLR + rcRLx * uRL + rcRRx * uRR +
                     (PtotLL - PtotRL + PtotLR - PtotRR),
                 rcx, recip(rcx));
  R vstar = divr(rcLLy * vLL + rcLRy * vLR + rcRLy * vRL + rcRRy * vRR +
                     (PtotLL - PtotLR + PtotRL - PtotRR),
                 rcy, recip(rcy));

  const R iSLu = recip(SL - ustar), iSRu = recip(SR - ustar);
  const R iSBv = recip(SB - vstar), iSTv = recip(ST - vstar);

  R rstarLLx = divr(rLL * (SL - uLL), SL - ustar, iSLu);
  R BstarLL = divr(bLL * (SL - uLL), SL - ustar, iSLu);
  R rstarLLy = divr(rLL * (SB - vLL), SB - vstar, iSBv);
  R AstarLL = divr(aLL * (SB - vLL), SB - vstar, iSBv);
  R rstarLL = divr(divr(rLL * (SL - uLL), SL - ustar, iSLu) * (SB - vLL), SB - vstar, iSBv);
  R EstarLLx = ustar * BstarLL - vLL * aLL;
  R EstarLLy = uLL * bLL - vstar * AstarLL;
  R EstarLL = ustar * BstarLL - vstar * AstarLL;

  R rstarLRx = divr(rLR * (SL - uLR), SL - ustar, iSLu);
  R BstarLR = divr(bLR * (SL - uLR), SL - ustar, iSLu);
  R rstarLRy = divr(rLR * (ST - vLR), ST - vstar, iSTv);
  R AstarLR = divr(aLR * (ST - vLR), ST - vstar, iSTv);
  R rstarLR = divr(divr(rLR * (SL - uLR), SL - ustar, iSLu) * (ST - vLR), ST - vstar, iSTv);
  R EstarLRx = ustar * BstarLR - vLR * aLR;
  R EstarLRy = uLR * bLR - vstar * AstarLR;
  R EstarLR = ustar * BstarLR - vstar * AstarLR;

  R rstarRLx = divr(rRL * (SR - uRL), SR - ustar, iSRu);
  R BstarRL = divr(bRL * (SR - uRL), SR - ustar, iSRu);
  R rstarRLy = divr(rRL * (SB - vRL), SB - vstar, iSBv);
  R AstarRL = divr(aRL * (SB - vRL), SB - vstar, iSBv);
  R rstarRL = divr(divr(rRL * (SR - uRL), SR - ustar, iSRu) * (SB - vRL), SB - vstar, iSBv);
  R EstarRLx = ustar * BstarRL - vRL * aRL;
  R EstarRLy = uRL * bRL - vstar * AstarRL;
  R EstarRL = ustar * BstarRL - vstar * AstarRL;

  R rstarRRx = divr(rRR * (SR - uRR), SR - ustar, iSRu);
  R BstarRR = divr(bRR * (SR - uRR), SR - ustar, iSRu);
  R rstarRRy = divr(rRR * (ST - vRR), ST - vstar, iSTv);
  R AstarRR = divr(aRR * (ST - vRR), ST - vstar, iSTv);
  R rstarRR = divr(divr(rRR * (SR - uRR), SR - ustar, iSRu) * (ST - vRR), ST - vstar, iSTv);
  R EstarRRx = ustar * BstarRR - vRR * aRR;
  R EstarRRy = uRR * bRR - vstar * AstarRR;
  R EstarRR = ustar * BstarRR - vstar * AstarRR;

  const R smallc = R(p.smallc);
  R calfvenL = max5(over_sqrt(dabs(aLR), rstarLRx), over_sqrt(dabs(AstarLR), rstarLR),
                    over_sqrt(dabs(aLL), rstarLLx), over_sqrt(dabs(AstarLL), rstarLL), smallc);
  R calfvenR = max5(over_sqrt(dabs(aRR), rstarRRx), over_sqrt(dabs(AstarRR), rstarRR),
                    over_sqrt(dabs(aRL), rstarRLx), over_sqrt(dabs(AstarRL), rstarRL), smallc);
  R calfvenB = max5(over_sqrt(dabs(bLL), rstarLLy), over_sqrt(dabs(BstarLL), rstarLL),
                    over_sqrt(dabs(bRL), rstarRLy), over_sqrt(dabs(BstarRL), rstarRL), smallc);
  R calfvenT = max5(over_sqrt(dabs(bLR), rstarLRy), over_sqrt(dabs(BstarLR), rstarLR),
                    over_sqrt(dabs(bRR), rstarRRy), over_sqrt(dabs(BstarRR), rstarRR), smallc);

  R SAL = dmin(ustar - calfvenL, R(0));
  R SAR = dmax(ustar + calfvenR, R(0));
  R SAB = dmin(vstar - calfvenB, R(0));
  R SAT = dmax(vstar + calfvenT, R(0));

  const R iSAx = recip(SAR - SAL), iSAy = recip(SAT - SAB);
  R AstarT = divr(SAR * AstarRR - SAL * AstarLR, SAR - SAL, iSAx);
  R AstarB = divr(SAR * AstarRL - SAL * AstarLL, SAR - SAL, iSAx);
  R BstarR = divr(SAT * BstarRR - SAB * BstarRL, SAT - SAB, iSAy);
  R BstarL = divr(SAT * BstarLR - SAB * BstarLL, SAT - SAB, iSAy);

  // the five region EMFs: strongly interacting (central), bottom, top, left, right
  const R Ec = divr(divr(SAL * SAB * EstarRR - SAL * SAT * EstarRL - SAR * SAB * EstarLR +
                             SAR * SAT * EstarLL,
                         SAR - SAL, iSAx),
                    SAT - SAB, iSAy) -
               divr(SAT * SAB, SAT - SAB, iSAy) * (AstarT - AstarB) +
               divr(SAR * SAL, SAR - SAL, iSAx) * (BstarR - BstarL);
  const R Eb = divr(SAR * EstarLLx - SAL * EstarRLx + SAR * SAL * (bRL - bLL), SAR - SAL, iSAx);
  const R Et = divr(SAR * EstarLRx - SAL * EstarRRx + SAR * SAL * (bRR - bLR), SAR - SAL, iSAx);
  const R El = divr(SAT * EstarLLy - SAB * EstarLRy - SAT * SAB * (aLR - aLL), SAT - SAB, iSAy);
  const R Er = divr(SAT * EstarRLy - SAB * EstarRRy - SAT * SAB * (aRR - aRL), SAT - SAB, iSAy);

  if (kSharedRecip) {
    // the default builds select the one region whose mask product is 1 (the masks partition
    // the sign cases; the blend below adds that term to zeros, which gives the same value for
    // finite region EMFs, the sign of a zero result aside)
    const bool sbp = !sign_bit(SB), stp = !sign_bit(ST), slp = !sign_bit(SL), srp = !sign_bit(SR);
    if (sbp) return slp ? ELL : (srp ? Eb : ERL);
    if (!stp) return slp ? ELR : (srp ? Et : ERR);
    if (slp) return El;
    return srp ? Ec : Er;
  }
  // 0/1 region masks from the sign bit (copysign(1, S) > 0), blended as the reference does
  const R SB_pos = sign_bit(SB) ? R(0) : R(1), SB_neg = R(1) - SB_pos;
  const R ST_pos = sign_bit(ST) ? R(0) : R(1), ST_neg = R(1) - ST_pos;
  const R SL_pos = sign_bit(SL) ? R(0) : R(1), SL_neg = R(1) - SL_pos;
  const R SR_pos = sign_bit(SR) ? R(0) : R(1), SR_neg = R(1) - SR_pos;

  R E = R(0);
  E += (SB_neg * ST_pos * SL_neg * SR_pos) * Ec;
  E += SB_pos * (SL_pos * ELL + SL_neg * SR_neg * ERL + SL_neg * SR_pos * Eb);
  E += (SB_neg * ST_neg) * (SL_pos * ELR + SL_neg * SR_neg * ERR + SL_neg * SR_pos * Et);
  E += (SB_neg * ST_pos * SL_pos) * El;
  E += (SB_neg * ST_pos * SL_neg * SR_neg) * Er;
  return E;
}

// compute_emf<emfDir>: map the 4 edge states (RT, RB, LT, LB) to LL/RL/LR/RR with the
// parallel field averaged, then the 2D HLLD solver, reference
// src/shared/RiemannSolvers_MHD.h:651-874.  Template parameter = (iu, iv, iw, ia, ib, ic)
// permutation: EMFZ (u,v,w), EMFY (w,u,v), EMFX (v,w,u).
template <int EDIR, class R>
PPK_HD R compute_emf(const St<R> &qRT, const St<R> &qRB, const St<R> &qLT,
                                         const St<R> &qLB, const DevParams &p) {
  constexpr int iu = EDIR == EMFZ ? IU : (EDIR == EMFY ? IW : IV);
  constexpr int iv = EDIR == EMFZ ? IV : (EDIR == EMFY ? IU : IW);
  constexpr int iw = EDIR == EMFZ ? IW : (EDIR == EMFY ? IV : IU);
  constexpr int ia = EDIR == EMFZ ? IA : (EDIR == EMFY ? IC : IB);
  constexpr int ib = EDIR == EMFZ ? IB : (EDIR == EMFY ? IA : IC);
  constexpr int ic = EDIR == EMFZ ? IC : (EDIR == EMFY ? IB : IA);
  St<R> qLL, qRL, qLR, qRR;
  qLL[ID] = qRT[ID]; qRL[ID] = qLT[ID]; qLR[ID] = qRB[ID]; qRR[ID] = qLB[ID];
  if (p.cIso > 0) {
    qLL[IP] = qLL[ID] * R(p.cIso) * R(p.cIso);
    qRL[IP] = qRL[ID] * R(p.cIso) * R(p.cIso);
    qLR[IP] = qLR[ID] * R(p.cIso) * R(p.cIso);
    qRR[IP] = qRR[ID] * R(p.cIso) * R(p.cIso);
  } else {
    qLL[IP] = qRT[IP]; qRL[IP] = qLT[IP]; qLR[IP] = qRB[IP]; qRR[IP] = qLB[IP];
  }
  qLL[IU] = qRT[iu]; qRL[IU] = qLT[iu]; qLR[IU] = qRB[iu]; qRR[IU] = qLB[iu];
  qLL[IV] = qRT[iv]; qRL[IV] = qLT[iv]; qLR[IV] = qRB[iv]; qRR[IV] = qLB[iv];
  const R ax_top = R(0.5) * (qRT[ia] + qLT[ia]);
  const R ax_bot = R(0.5) * (qRB[ia] + qLB[ia]);
  qLL[IA] = ax_top; qRL[IA] = ax_top; qLR[IA] = ax_bot; qRR[IA] = ax_bot;
  const R by_right = R(0.5) * (qRT[ib] + qRB[ib]);
  const R by_left = R(0.5) * (qLT[ib] + qLB[ib]);
  qLL[IB] = by_right; qRL[IB] = by_left; qLR[IB] = by_right; qRR[IB] = by_left;
  qLL[IW] = qRT[iw]; qRL[IW] = qLT[iw]; qLR[IW] = qRB[iw]; qRR[IW] = qLB[iw];
  qLL[IC] = qRT[ic]; qRL[IC] = qLT[ic]; qLR[IC] = qRB[ic]; qRR[IC] = qLB[ic];
  R ELL = qLL[IU] * qLL[IB] - qLL[IV] * qLL[IA];
  R ERL = qRL[IU] * qRL[IB] - qRL[IV] * qRL[IA];
  R ELR = qLR[IU] * qLR[IB] - qLR[IV] * qLR[IA];
  R ERR = qRR[IU] * qRR[IB] - qRR[IV] * qRR[IA];
  return mag_riemann2d_hlld(qLL, qRL, qLR, qRR, ELL, ERL, ELR, ERR, p);
}

}  // namespace ppk
