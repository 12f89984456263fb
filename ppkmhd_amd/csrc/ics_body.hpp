// Initial conditions of the MHD problems, cell by cell: host+device functions shared by the
// host initialisation (ics.cpp) and the on-device one (ics_device.hip), so both run the same
// operations in the same order.  Per-problem constants that involve transcendental functions
// of the parameters only (B0, the wave angles, ...) are computed once on the host
// (ic_consts) and passed to both; per-cell sin / cos / exp come from the host or the device
// math library.  Reference functors: src/muscl/MHDInitFunctors{2D,3D}.h.
#pragma once

#include <cstdint>

#include "mhd_device.hpp"
#include "ppkmhd_amd.h"

namespace ppk {
namespace icb {

struct IcBox {
  int is, js, ks, gw;
  PPK_HD long long at(int i, int j, int k, int v) const {
    return (((long long)v * ks + k) * js + j) * (long long)is + i;
  }
  PPK_HD long long cells() const { return (long long)is * js * ks; }
};
inline IcBox box_of(const ppk_mhd_params &p) {
  IcBox b;
  b.gw = p.ghost_width;
  b.is = p.nx + 2 * b.gw;
  b.js = p.ny + 2 * b.gw;
  b.ks = p.dim == 3 ? p.nz + 2 * b.gw : 1;
  return b;
}

// constants of a problem that depend on its parameters only
struct IcConsts {
  double B0, p0, d0, p0e;  // Orszag-Tang (p0e: the energy's p0, from 4 asin(1))
  // linear wave (WaveParams, src/shared/problems/WaveParams.h:31-146)
  double rev[7], V0, amp, wd0, wp0, cos_a2, sin_a2, cos_a3, sin_a3, k_par, dby, dbz, bx0,
      by0, bz0;
};

constexpr double kPi = 3.141592653589793238462643383279502884;

// counter-based generator for the random KH layer: splitmix64 of (seed, global cell, draw).
// The reference draws from a Kokkos XorShift64 pool whose stream depends on the thread
// mapping, so no generator reproduces it; this one is identical on every device.
PPK_HD double kh_uniform(uint64_t seed, uint64_t cell, uint64_t draw) {
  uint64_t x = seed * 0x9E3779B97F4A7C15ull + cell * 0xD1B54A32D192ED03ull +
               draw * 0x8CB92BA72F3D8DD7ull;
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  x = x ^ (x >> 31);
  return (double)(x >> 11) * (1.0 / 9007199254740992.0);
}

// Every problem except the linear wave: all variables of cell (i,j,k), except the total
// energy of Orszag-Tang (it needs the neighbours' face fields: ic_ot_energy).  Every cell of
// the padded box, as the reference's init functors run over the full range.
PPK_HD void ic_cell(const ppk_mhd_params &p, const ppk_mhd_problem &pr, const IcConsts &K,
                    const IcBox &b, int i, int j, int k, double *U) {
  const int gw = b.gw;
  const double gamma0 = p.gamma0;
  const double TWOPI = 2 * kPi;
  auto at = [&](int v) -> double & { return U[b.at(i, j, k, v)]; };
  const double x = p.xmin + p.dx / 2 + (i - gw) * p.dx;
  const double y = p.ymin + p.dy / 2 + (j - gw) * p.dy;
  const double z = p.dim == 3 ? p.zmin + p.dz / 2 + (k - gw) * p.dz : 0.0;
  if (pr.kind == PPK_PROBLEM_ORSZAG_TANG) {
    // InitOrszagTangFunctor2D / 3D, src/muscl/MHDInitFunctors2D.h:289-383,
    // src/muscl/MHDInitFunctors3D.h:311-409
    const double v0 = 1.0;
    at(ID) = K.d0;
    at(IU) = -K.d0 * v0 * sin(y * TWOPI);
    at(IV) = K.d0 * v0 * sin(x * TWOPI);
    at(IW) = 0.0;
    if (p.dim == 2) {
      at(IA) = -K.B0 * sin(y * TWOPI);
      at(IB) = K.B0 * sin(2.0 * x * TWOPI);
    } else {
      const double cz = cos(2 * TWOPI * pr.ot_kt * (z - p.zmin) / (p.zmax - p.zmin));
      at(IA) = -K.B0 * cz * sin(y * TWOPI);
      at(IB) = K.B0 * cz * sin(2.0 * x * TWOPI);
    }
    at(IC) = 0.0;
  } else if (pr.kind == PPK_PROBLEM_BLAST) {
    // InitBlastFunctor{2D,3D}_MHD, src/muscl/MHDInitFunctors2D.h:164-231,
    // src/muscl/MHDInitFunctors3D.h:175-254 (uniform B is 0.5 in the reference)
    const double r2 = pr.blast_radius * pr.blast_radius;
    double d2;
    if (p.dim == 3)
      d2 = (x - pr.blast_center_x) * (x - pr.blast_center_x) +
           (y - pr.blast_center_y) * (y - pr.blast_center_y) +
           (z - pr.blast_center_z) * (z - pr.blast_center_z);
    else
      d2 = (x - pr.blast_center_x) * (x - pr.blast_center_x) +
           (y - pr.blast_center_y) * (y - pr.blast_center_y);
    const bool in = d2 < r2;
    at(ID) = in ? pr.blast_density_in : pr.blast_density_out;
    at(IU) = 0.0;
    at(IV) = 0.0;
    at(IW) = 0.0;
    at(IA) = pr.blast_bx;
    at(IB) = pr.blast_by;
    at(IC) = pr.blast_bz;
    const double a = pr.blast_bx, bb = pr.blast_by, c = pr.blast_bz;
    at(IP) = (in ? pr.blast_pressure_in : pr.blast_pressure_out) / (gamma0 - 1.0) +
             0.5 * (a * a + bb * bb + c * c);
  } else if (pr.kind == PPK_PROBLEM_KELVIN_HELMHOLTZ) {
    // InitKelvinHelmholtzFunctor{2D,3D}_MHD, src/muscl/MHDInitFunctors2D.h:416-565,
    // src/muscl/MHDInitFunctors3D.h:441-604
    const double d_in = pr.kh_d_in, d_out = pr.kh_d_out, vin = pr.kh_vflow_in;
    const double vout = pr.kh_vflow_out, ampl = pr.kh_amplitude, pressure = pr.kh_pressure;
    const uint64_t cell = ((uint64_t)k * b.js + j) * b.is + i;
    if (p.dim == 2) {
      if (pr.kh_p_rand) {
        const double yn = (y - p.ymin) / (p.ymax - p.ymin);
        double d, u, v;
        if (yn < 0.25 || yn > 0.75) { d = d_out; u = vout; v = 0.0; }
        else { d = d_in; u = vin; v = 0.0; }
        u += ampl * (kh_uniform(pr.kh_seed, cell, 0) - 0.5);
        v += ampl * (kh_uniform(pr.kh_seed, cell, 1) - 0.5);
        const double bx = 0.5, by = 0.0, bz = 0.0;
        at(ID) = d; at(IU) = d * u; at(IV) = d * v;
        at(IA) = bx; at(IB) = by; at(IC) = bz;
        at(IP) = pressure / (gamma0 - 1.0) + 0.5 * d * (u * u + v * v) +
                 0.5 * (bx * bx + by * by + bz * bz);
      } else if (pr.kh_p_sine_rob || pr.kh_p_sine) {
        const int n = pr.kh_mode;
        const double w0 = pr.kh_w0, delta = pr.kh_delta, y1 = 0.25, y2 = 0.75;
        double d, u;
        if (pr.kh_p_sine_rob) {
          const double ramp = 1.0 / (1.0 + exp(2 * (y - y1) / delta)) +
                              1.0 / (1.0 + exp(2 * (y2 - y) / delta));
          d = d_in + ramp * (d_out - d_in);
          u = vin + ramp * (vout - vin);
        } else {
          d = (y >= y1 && y <= y2) ? d_in : d_out;
          u = (y >= y1 && y <= y2) ? vin : vout;
        }
        const double v = w0 * sin(n * kPi * x);
        const double bx = 0.5, by = 0.0, bz = 0.0;
        at(ID) = d; at(IU) = d * u; at(IV) = d * v;
        at(IA) = bx; at(IB) = by; at(IC) = bz;
        at(IP) = pressure / (gamma0 - 1.0) + 0.5 * d * (u * u + v * v) +
                 0.5 * (bx * bx + by * by + bz * bz);
      }
    } else {
      // shear-normal coordinate: z in the reference; y when kh_axis == 1 (extension)
      const bool yaxis = pr.kh_axis == 1;
      const double s = yaxis ? y : z;
      const double smin = yaxis ? p.ymin : p.zmin, smax = yaxis ? p.ymax : p.zmax;
      if (pr.kh_p_rand) {
        const double sn = (s - smin) / (smax - smin);
        double d, u, v, w;
        if (sn < 0.25 || sn > 0.75) { d = d_out; u = vout; v = 0.0; w = 0.0; }
        else { d = d_in; u = vin; v = 0.0; w = 0.0; }
        u += ampl * (kh_uniform(pr.kh_seed, cell, 0) - 0.5);
        v += ampl * (kh_uniform(pr.kh_seed, cell, 1) - 0.5);
        if (!yaxis) w += ampl * (kh_uniform(pr.kh_seed, cell, 2) - 0.5);
        const double bx = pr.kh_bx, by = pr.kh_by, bz = pr.kh_bz;
        at(ID) = d; at(IU) = d * u; at(IV) = d * v; at(IW) = d * w;
        at(IA) = bx; at(IB) = by; at(IC) = bz;
        at(IP) = pressure / (gamma0 - 1.0) + 0.5 * d * (u * u + v * v + w * w) +
                 0.5 * (bx * bx + by * by + bz * bz);
      } else if (pr.kh_p_sine_rob) {
        const int n = pr.kh_mode;
        const double w0 = pr.kh_w0, delta = pr.kh_delta, z1 = 0.25, z2 = 0.75;
        const double ramp = 1.0 / (1.0 + exp(2 * (z - z1) / delta)) +
                            1.0 / (1.0 + exp(2 * (z2 - z) / delta));
        const double v1y = vin / 2, v2y = vout / 2;
        const double d = d_in + ramp * (d_out - d_in);
        const double u = vin + ramp * (vout - vin);
        const double v = v1y + ramp * (v2y - v1y);
        const double w = w0 * sin(n * kPi * x) * sin(n * kPi * y);
        const double bx = 0.5, by = 0.0, bz = 0.0;
        at(ID) = d; at(IU) = d * u; at(IV) = d * v; at(IW) = d * w;
        at(IA) = bx; at(IB) = by; at(IC) = bz;
        at(IP) = pressure / (gamma0 - 1.0) + 0.5 * d * (u * u + v * v + w * w) +
                 0.5 * (bx * bx + by * by + bz * bz);
      } else if (pr.kh_p_sine) {
        const int n = pr.kh_mode;
        const double w0 = pr.kh_w0, z1 = 0.25, z2 = 0.75;
        const double d = (z >= z1 && z <= z2) ? d_in : d_out;
        const double u = (z >= z1 && z <= z2) ? vin : vout;
        const double v = 0;
        const double w = w0 * sin(n * kPi * x);
        const double bx = 0.5, by = 0.0, bz = 0.0;
        at(ID) = d; at(IU) = d * u; at(IV) = d * v; at(IW) = d * w;
        at(IA) = bx; at(IB) = by; at(IC) = bz;
        at(IP) = pressure / (gamma0 - 1.0) + 0.5 * d * (u * u + v * v + w * w) +
                 0.5 * (bx * bx + by * by + bz * bz);
      }
    }
  } else if (pr.kind == PPK_PROBLEM_IMPLODE) {
    // InitImplodeFunctor{2D,3D}_MHD, src/muscl/MHDInitFunctors2D.h:58-134,
    // src/muscl/MHDInitFunctors3D.h:56-145
    const double *in = pr.implode_in, *out = pr.implode_out;
    bool tmp;
    if (p.dim == 2) {
      if (pr.implode_shape == 1) tmp = x + y > 0.5 && x + y < 2.5;
      else tmp = x + y > (p.xmin + p.xmax) / 2. + p.ymin;
    } else {
      if (pr.implode_shape == 1) tmp = x + y + z > 0.5 && x + y + z < 2.5;
      else tmp = x + y + z > (p.xmin + p.xmax) / 2. + p.ymin + p.zmin;
    }
    const double *s = tmp ? out : in;
    at(ID) = s[0];
    if (p.dim == 2)
      at(IP) = s[1] / (gamma0 - 1.0) + 0.5 * s[0] * (s[2] * s[2] + s[3] * s[3]) +
               0.5 * (s[5] * s[5] + s[6] * s[6]);
    else
      at(IP) = s[1] / (gamma0 - 1.0) + 0.5 * s[0] * (s[2] * s[2] + s[3] * s[3] + s[4] * s[4]) +
               0.5 * (s[5] * s[5] + s[6] * s[6] + s[7] * s[7]);
    at(IU) = s[2];
    at(IV) = s[3];
    at(IW) = p.dim == 2 ? 0.0 : s[4];
    at(IA) = s[5];
    at(IB) = s[6];
    at(IC) = p.dim == 2 ? 0.0 : s[7];
  } else if (pr.kind == PPK_PROBLEM_ROTOR) {
    // InitRotorFunctor2D_MHD / 3D_MHD, src/muscl/MHDInitFunctors2D.h:588-712,
    // src/muscl/MHDInitFunctors3D.h:627-756: a dense spinning disk (radius r0, a linear ramp to
    // r1) in a uniform field b0 along x.  The 2D functor writes momenta with the ramp factor
    // f_r also inside r0; the 3D one writes velocities in the momentum slots; both kept.
    const double xc = (p.xmax + p.xmin) / 2, yc = (p.ymax + p.ymin) / 2;
    const double r0 = pr.rotor_r0, r1 = pr.rotor_r1, u0 = pr.rotor_u0;
    const double r = sqrt((x - xc) * (x - xc) + (y - yc) * (y - yc));
    const double f_r = (r1 - r) / (r1 - r0);
    double d, mu, mv;
    if (r <= r0) {
      d = 10.0;
      mu = p.dim == 2 ? -d * f_r * u0 * (y - yc) / r0 : -u0 * (y - yc) / r0;
      mv = p.dim == 2 ? d * f_r * u0 * (x - xc) / r0 : u0 * (x - xc) / r0;
    } else if (r <= r1) {
      d = 1 + 9 * f_r;
      mu = p.dim == 2 ? -d * f_r * u0 * (y - yc) / r : -f_r * u0 * (y - yc) / r;
      mv = p.dim == 2 ? d * f_r * u0 * (x - xc) / r : f_r * u0 * (x - xc) / r;
    } else {
      d = 1.0;
      mu = 0.0;
      mv = 0.0;
    }
    const double mw = 0.0, a = pr.rotor_b0;
    at(ID) = d;
    at(IU) = mu;
    at(IV) = mv;
    at(IW) = mw;
    at(IA) = a;
    at(IB) = 0.0;
    at(IC) = 0.0;
    at(IP) = p.dim == 2 ? pr.rotor_p0 / (gamma0 - 1.0) + 0.5 * (mu * mu + mv * mv + mw * mw) / d +
                              0.5 * (a * a)
                        : pr.rotor_p0 / (gamma0 - 1.0) + (mu * mu + mv * mv + mw * mw) / 2 / d +
                              (a * a) / 2;
  }
}

// Field loop (InitFieldLoopFunctor2D_MHD / 3D_MHD, src/muscl/MHDInitFunctors2D.h:715-893,
// src/muscl/MHDInitFunctors3D.h:759-1010), in three passes as the reference's functor: the
// vector potential A_z = amplitude (radius - r) inside the loop (r from the origin, every
// cell), ...
PPK_HD void ic_loop_potential(const ppk_mhd_params &p, const ppk_mhd_problem &pr,
                              const IcBox &b, int i, int j, int k, double *A) {
  const double x = p.xmin + p.dx / 2 + (i - b.gw) * p.dx;
  const double y = p.ymin + p.dy / 2 + (j - b.gw) * p.dy;
  const double r = sqrt(x * x + y * y);
  A[b.at(i, j, k, IX)] = 0.0;
  A[b.at(i, j, k, IY)] = 0.0;
  A[b.at(i, j, k, IZ)] = r < pr.loop_radius ? pr.loop_amplitude * (pr.loop_radius - r) : 0.0;
}
// ... the density, the flow along the diagonal and the face fields as the discrete curl of A,
// interior cells (the 2D functor gives the flow a z component rho vflow nz / |(nx, ny, nz)|
// that its energy leaves out) ...
PPK_HD void ic_loop_state(const ppk_mhd_params &p, const ppk_mhd_problem &pr, const IcBox &b,
                          int i, int j, int k, const double *A, double *U) {
  auto Aat = [&](int ii, int jj, int kk, int c) { return A[b.at(ii, jj, kk, c)]; };
  auto at = [&](int v) -> double & { return U[b.at(i, j, k, v)]; };
  const double x = p.xmin + p.dx / 2 + (i - b.gw) * p.dx;
  const double y = p.ymin + p.dy / 2 + (j - b.gw) * p.dy;
  const double cos_theta = 2.0 / sqrt(5.0);
  const double sin_theta = sqrt(1 - cos_theta * cos_theta);
  const double r = sqrt(x * x + y * y);
  const double d = r < pr.loop_radius ? pr.loop_density_in : 1.0;
  const double vflow = pr.loop_vflow;
  at(ID) = d;
  at(IU) = d * vflow * cos_theta;
  at(IV) = d * vflow * sin_theta;
  if (p.dim == 2) {
    const double diag = sqrt(1.0 * (p.nx * p.nx + p.ny * p.ny + p.nz * p.nz));
    at(IW) = d * vflow * p.nz / diag;
    at(IA) = (Aat(i, j + 1, k, IZ) - Aat(i, j, k, IZ)) / p.dy;
    at(IB) = -(Aat(i + 1, j, k, IZ) - Aat(i, j, k, IZ)) / p.dx;
    at(IC) = 0.0;
  } else {
    at(IW) = d * vflow;
    at(IA) = (Aat(i, j + 1, k, 2) - Aat(i, j, k, 2)) / p.dy -
             (Aat(i, j, k + 1, 1) - Aat(i, j, k, 1)) / p.dz;
    at(IB) = (Aat(i, j, k + 1, 0) - Aat(i, j, k, 0)) / p.dz -
             (Aat(i + 1, j, k, 2) - Aat(i, j, k, 2)) / p.dx;
    at(IC) = (Aat(i + 1, j, k, 1) - Aat(i, j, k, 1)) / p.dx -
             (Aat(i, j + 1, k, 0) - Aat(i, j, k, 0)) / p.dy;
  }
}
// ... and the total energy of interior cells once every face field is set (the last interior
// column / row averages with the +1 ghost face, still zero: the reference fills ghosts after)
PPK_HD void ic_loop_energy(const ppk_mhd_params &p, const IcBox &b, int i, int j, int k,
                           double *U) {
  auto at = [&](int ii, int jj, int kk, int v) -> double & { return U[b.at(ii, jj, kk, v)]; };
  if (p.cIso > 0) {
    at(i, j, k, IP) = 0.0;
    return;
  }
  const double gamma0 = p.gamma0;
  const double sa = at(i, j, k, IA) + at(i + 1, j, k, IA);
  const double sb = at(i, j, k, IB) + at(i, j + 1, k, IB);
  const double u = at(i, j, k, IU), v = at(i, j, k, IV), w = at(i, j, k, IW);
  if (p.dim == 2) {
    at(i, j, k, IP) = 1.0f / (gamma0 - 1.0) + 0.5 * (0.25 * (sa * sa) + 0.25 * (sb * sb)) +
                      0.5 * (u * u + v * v) / at(i, j, k, ID);
  } else {
    const double sc = at(i, j, k, IC) + at(i, j, k + 1, IC);
    at(i, j, k, IP) = 1.0f / (gamma0 - 1.0) +
                      0.5 * (0.25 * (sa * sa) + 0.25 * (sb * sb) + 0.25 * (sc * sc)) +
                      0.5 * (u * u + v * v + w * w) / at(i, j, k, ID);
  }
}

// Orszag-Tang total energy of cell (i,j,k), i < is - 1, j < js - 1 (the reference's second
// loop): kinetic energy and the energy of the cell-centred field from the two faces
PPK_HD void ic_ot_energy(const ppk_mhd_params &p, const IcConsts &K, const IcBox &b, int i,
                         int j, int k, double *U) {
  auto at = [&](int ii, int jj, int v) -> double & { return U[b.at(ii, jj, k, v)]; };
  const double gamma0 = p.gamma0;
  const double d = at(i, j, ID);
  const double pe = p.dim == 2 ? K.p0e : K.p0;
  at(i, j, IP) = pe / (gamma0 - 1.0) +
                 0.5 * (at(i, j, IU) * at(i, j, IU) / d + at(i, j, IV) * at(i, j, IV) / d +
                        0.25 * ((at(i, j, IA) + at(i + 1, j, IA)) *
                                (at(i, j, IA) + at(i + 1, j, IA))) +
                        0.25 * ((at(i, j, IB) + at(i, j + 1, IB)) *
                                (at(i, j, IB) + at(i, j + 1, IB))));
}

// linear wave, 3D (InitWaveFunctor3D_MHD, src/muscl/MHDInitFunctors3D.h:1098-1298): the
// vector potential at the cell's edges (A has 3 components over the padded box) ...
PPK_HD void ic_wave_potential(const ppk_mhd_params &p, const IcConsts &K, const IcBox &b, int i,
                              int j, int k, double *A) {
  const int gw = b.gw;
  const double x = p.xmin + p.dx / 2 + (i - gw) * p.dx;
  const double y = p.ymin + p.dy / 2 + (j - gw) * p.dy;
  const double z = p.zmin + p.dz / 2 + (k - gw) * p.dz;
  double Ay, Az, tmpx, tmpy, X1, X2, X3;
  X1 = x; X2 = y - p.dy / 2; X3 = z - p.dz / 2;
  tmpx = X1 * K.cos_a2 * K.cos_a3 + X2 * K.cos_a2 * K.sin_a3 + X3 * K.sin_a2;
  tmpy = -X1 * K.sin_a3 + X2 * K.cos_a3;
  Ay = K.bz0 * tmpx - (K.dbz / K.k_par) * cos(K.k_par * tmpx);
  Az = -K.by0 * tmpx + (K.dby / K.k_par) * cos(K.k_par * tmpx) + K.bx0 * tmpy;
  A[b.at(i, j, k, IX)] = -Ay * K.sin_a3 - Az * K.sin_a2 * K.cos_a3;
  X1 = x - p.dx / 2; X2 = y; X3 = z - p.dz / 2;
  tmpx = X1 * K.cos_a2 * K.cos_a3 + X2 * K.cos_a2 * K.sin_a3 + X3 * K.sin_a2;
  tmpy = -X1 * K.sin_a3 + X2 * K.cos_a3;
  Ay = K.bz0 * tmpx - (K.dbz / K.k_par) * cos(K.k_par * tmpx);
  Az = -K.by0 * tmpx + (K.dby / K.k_par) * cos(K.k_par * tmpx) + K.bx0 * tmpy;
  A[b.at(i, j, k, IY)] = Ay * K.cos_a3 - Az * K.sin_a2 * K.sin_a3;
  X1 = x - p.dx / 2; X2 = y - p.dy / 2; X3 = z;
  tmpx = X1 * K.cos_a2 * K.cos_a3 + X2 * K.cos_a2 * K.sin_a3 + X3 * K.sin_a2;
  tmpy = -X1 * K.sin_a3 + X2 * K.cos_a3;
  Az = -K.by0 * tmpx + (K.dby / K.k_par) * cos(K.k_par * tmpx) + K.bx0 * tmpy;
  A[b.at(i, j, k, IZ)] = Az * K.cos_a2;
}
// ... the face fields as its discrete curl, cells [gw-1, size-gw+1) ...
PPK_HD void ic_wave_field(const ppk_mhd_params &p, const IcBox &b, int i, int j, int k,
                          const double *A, double *U) {
  auto Aat = [&](int ii, int jj, int kk, int c) { return A[b.at(ii, jj, kk, c)]; };
  U[b.at(i, j, k, IA)] = (Aat(i, j + 1, k, 2) - Aat(i, j, k, 2)) / p.dy -
                         (Aat(i, j, k + 1, 1) - Aat(i, j, k, 1)) / p.dz;
  U[b.at(i, j, k, IB)] = (Aat(i, j, k + 1, 0) - Aat(i, j, k, 0)) / p.dz -
                         (Aat(i + 1, j, k, 2) - Aat(i, j, k, 2)) / p.dx;
  U[b.at(i, j, k, IC)] = (Aat(i + 1, j, k, 1) - Aat(i, j, k, 1)) / p.dx -
                         (Aat(i, j + 1, k, 0) - Aat(i, j, k, 0)) / p.dy;
}
// ... and the fluid state, interior cells
PPK_HD void ic_wave_state(const ppk_mhd_params &p, const IcConsts &K, const IcBox &b, int i,
                          int j, int k, double *U) {
  const int gw = b.gw;
  const double gamma0 = p.gamma0;
  const double x = p.xmin + p.dx / 2 + (i - gw) * p.dx;
  const double y = p.ymin + p.dy / 2 + (j - gw) * p.dy;
  const double z = p.zmin + p.dz / 2 + (k - gw) * p.dz;
  const double X = K.cos_a2 * (x * K.cos_a3 + y * K.sin_a3) + z * K.sin_a2;
  const double sn = sin(K.k_par * X);
  const double Mx = K.wd0 * K.V0 + K.amp * sn * K.rev[1];
  const double My = K.amp * sn * K.rev[2];
  const double Mz = K.amp * sn * K.rev[3];
  U[b.at(i, j, k, ID)] = K.wd0 + K.amp * sn * K.rev[0];
  U[b.at(i, j, k, IU)] = Mx * K.cos_a2 * K.cos_a3 - My * K.sin_a3 - Mz * K.sin_a2 * K.cos_a3;
  U[b.at(i, j, k, IV)] = Mx * K.cos_a2 * K.sin_a3 + My * K.cos_a3 - Mz * K.sin_a2 * K.sin_a3;
  U[b.at(i, j, k, IW)] = Mx * K.sin_a2 + Mz * K.cos_a2;
  U[b.at(i, j, k, IP)] = K.wp0 / (gamma0 - 1.0) + 0.5 * K.wd0 * K.V0 * K.V0 +
                         0.5 * (K.bx0 * K.bx0 + K.by0 * K.by0 + K.bz0 * K.bz0) +
                         K.amp * sn * K.rev[4];
}

}  // namespace icb

// the constants of a problem (host: libm)
icb::IcConsts ic_consts(const ppk_mhd_params &p, const ppk_mhd_problem &pr);

}  // namespace ppk
