// Learned HLLD flux (the surrogate form of the face solve): per-face pieces shared by the
// device kernels and host builds.
//
// At a face, in the frame rotated so that the normal is x (as the flux functors rotate,
// src/muscl/MHDRunFunctors3D.h:1839-1902), the surrogate flux is
//     F = F_LLF + scale * (mu_out + sig_out * MLP(standardised features))
// where F_LLF is the local Lax-Friedrichs flux of the two states (src/shared/
// RiemannSolvers_MHD.h:81-111), the MLP (16 -> 128 -> 128 -> 8, ReLU) has learned the
// difference F_HLLD - F_LLF, and `scale` makes that difference dimensionless.  The features
// use the similarity of ideal MHD under rho -> a rho, v -> b v, B -> sqrt(a) b B, p -> a b^2 p
// (fluxes scale as a b, a b^2, a b^3, sqrt(a) b^2 for mass, momentum, energy, field): with
// rho_bar = (rho_L + rho_R)/2 and s the LLF signal speed max(c_f + |u|) of the two states,
// the features are the dimensionless means and jumps of the two states,
//     x = [(rho_R - rho_L)/rho_bar, mean u, v, w / s, Bx / (sqrt(rho_bar) s), mean By, Bz / (same),
//          mean p / (rho_bar s^2), gamma - 5/3, jumps of u, v, w / s, of By, Bz / (sqrt(rho_bar) s),
//          of p / (rho_bar s^2), 0]
// (the same information as the two states: the mean density is the unit, Bx is one value
// after the solvers' averaging, and members of one ensemble differ in gamma).
//
// A face uses the surrogate only if it passes the gate (density and pressure jumps below
// thresholds) and its flux passes an a-posteriori admissibility check: the two states
// U_L - (F - F_L)/s and U_R + (F - F_R)/s, whose convex combinations make up the cell updates
// under the CFL condition, must have positive density and pressure (the property the LLF flux
// has by construction).  Any other face takes the exact HLLD flux.
#pragma once

#include "kernels.hpp"
#include "mhd_device.hpp"
#include "mlp.hpp"

namespace ppk {
// surrogate outputs: ID, IP, IU, IV, IW, IA (always 0), IB, IC -- the flux components in the
// reference's variable order
constexpr double kGammaRef = 5.0 / 3.0;

namespace PPK_NS {

// 1/x: in the default device build the shared-reciprocal forms of mhd_device.hpp (float: one
// v_rcp_f32, about 1 ulp, far below the network's own error; double: the refined reciprocal),
// the IEEE quotient in the strict build and on the host
template <class R>
PPK_HD R inv(R x) {
  return kSharedRecip ? recip(x) : R(1) / x;
}

// the LLF flux and its pieces (conservative states, physical fluxes, signal speed); same
// operations as riemann_llf
template <class R>
PPK_HD void llf_parts(St<R> &ql, St<R> &qr, St<R> &ul, St<R> &ur, St<R> &fl, St<R> &fr, R &s,
                      St<R> &F, const DevParams &par) {
  R bx_mean = R(0.5) * (ql[IA] + qr[IA]);
  ql[IA] = bx_mean;
  qr[IA] = bx_mean;
  mhd_flux(ql, ul, fl, par);
  mhd_flux(qr, ur, fr, par);
#pragma unroll
  for (int v = 0; v < 8; ++v) F[v] = (fl[v] + fr[v]) / R(2);
  auto info = [&](const St<R> &q) {
    R d = q[ID], pp = q[IP], a = q[IA], b = q[IB], c = q[IC];
    R b2 = a * a + b * b + c * c;
    const R id = inv(d);
    R c2 = divr(R(par.gamma0) * pp, d, id);
    R d2 = R(0.5) * (divr(b2, d, id) + c2);
    R cf = msqrt(d2 + msqrt(nonneg(d2 * d2 - divr(c2 * a * a, d, id))));
    return cf + dabs(q[IU]);
  };
  s = dmax(info(ql), info(qr));
#pragma unroll
  for (int v = 0; v < 8; ++v) F[v] = F[v] - s * (ur[v] - ul[v]) / R(2);
}

// dimensionless features of a face (after llf_parts: ql[IA] == qr[IA] == Bx): the means and
// the jumps of the two states, so that the small jumps of smooth faces are not lost to the
// bf16 rounding of the network inputs
template <class R>
PPK_HD void face_features(const St<R> &ql, const St<R> &qr, R s, R gamma0, float x[kMlpIn],
                          R &rho_bar) {
  rho_bar = R(0.5) * (ql[ID] + qr[ID]);
  const R iv = inv(s), ib = inv(msqrt(rho_bar) * s), ip = inv(rho_bar * s * s);
  const R ir = inv(rho_bar), h = R(0.5);
  x[0] = float((qr[ID] - ql[ID]) * ir);
  x[1] = float(h * (ql[IU] + qr[IU]) * iv);
  x[2] = float(h * (ql[IV] + qr[IV]) * iv);
  x[3] = float(h * (ql[IW] + qr[IW]) * iv);
  x[4] = float(ql[IA] * ib);
  x[5] = float(h * (ql[IB] + qr[IB]) * ib);
  x[6] = float(h * (ql[IC] + qr[IC]) * ib);
  x[7] = float(h * (ql[IP] + qr[IP]) * ip);
  x[8] = float(gamma0 - R(kGammaRef));
  x[9] = float((qr[IU] - ql[IU]) * iv);
  x[10] = float((qr[IV] - ql[IV]) * iv);
  x[11] = float((qr[IW] - ql[IW]) * iv);
  x[12] = float((qr[IB] - ql[IB]) * ib);
  x[13] = float((qr[IC] - ql[IC]) * ib);
  x[14] = float((qr[IP] - ql[IP]) * ip);
  x[15] = 0.f;
}

// physical scale of each output component
template <class R>
PPK_HD void out_scales(R rho_bar, R s, R sc[kMlpOut]) {
  const R m = rho_bar * s, b = msqrt(rho_bar) * s * s;
  sc[0] = m;
  sc[1] = m * s * s;
  sc[2] = m * s;
  sc[3] = m * s;
  sc[4] = m * s;
  sc[5] = R(0);
  sc[6] = b;
  sc[7] = b;
}

// density / pressure jump gate (strict: thresholds of 0 admit no face)
template <class R>
PPK_HD bool smooth_face(const St<R> &ql, const St<R> &qr, double gate_rho, double gate_p) {
  const R dr = dabs(qr[ID] - ql[ID]), dp = dabs(qr[IP] - ql[IP]);
  return dr < R(gate_rho) * dmin(ql[ID], qr[ID]) && dp < R(gate_p) * dmin(ql[IP], qr[IP]);
}

// pressure of a conservative state
template <class R>
PPK_HD R pressure_of(const St<R> &u, R gamma0) {
  const R d = u[ID];
  const R ke = divr(R(0.5) * (u[IU] * u[IU] + u[IV] * u[IV] + u[IW] * u[IW]), d, inv(d));
  const R me = R(0.5) * (u[IA] * u[IA] + u[IB] * u[IB] + u[IC] * u[IC]);
  return (gamma0 - R(1)) * (u[IP] - ke - me);
}

// a-posteriori check of a candidate flux F (see the file comment)
template <class R>
PPK_HD bool admissible(const St<R> &ul, const St<R> &ur, const St<R> &fl, const St<R> &fr,
                       const St<R> &F, R s, R gamma0) {
  St<R> a, b;
  const R is = inv(s);
#pragma unroll
  for (int v = 0; v < 8; ++v) {
    a[v] = ul[v] - (F[v] - fl[v]) * is;
    b[v] = ur[v] + (F[v] - fr[v]) * is;
  }
  return a[ID] > R(0) && b[ID] > R(0) && pressure_of(a, gamma0) > R(0) &&
         pressure_of(b, gamma0) > R(0);
}

// training target of a face: (F_HLLD - F_LLF) / scale, component by component (0 for IA)
template <class R>
PPK_HD void face_target(const St<R> &fh, const St<R> &F, const R sc[kMlpOut], float t[kMlpOut]) {
#pragma unroll
  for (int o = 0; o < kMlpOut; ++o) t[o] = o == IA ? 0.f : float((fh[o] - F[o]) / sc[o]);
}

}  // namespace PPK_NS
}  // namespace ppk
