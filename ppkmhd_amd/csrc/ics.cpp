// Initial conditions of the MHD problems, computed on the host in the reference layout
// (then uploaded).  Restates src/muscl/MHDInitFunctors2D.h and src/muscl/MHDInitFunctors3D.h
// (Orszag-Tang, blast, Kelvin-Helmholtz, linear wave, implode); citations per branch.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "ppkmhd_amd.h"

namespace ppk {
namespace {

enum { ID = 0, IP = 1, IU = 2, IV = 3, IW = 4, IA = 5, IB = 6, IC = 7 };
enum { IX = 0, IY = 1, IZ = 2 };

struct Box {
  int nx, ny, nz, gw, is, js, ks, dim;
  explicit Box(const ppk_mhd_params &p) {
    dim = p.dim;
    gw = p.ghost_width;
    nx = p.nx;
    ny = p.ny;
    nz = dim == 3 ? p.nz : 1;
    is = nx + 2 * gw;
    js = ny + 2 * gw;
    ks = dim == 3 ? nz + 2 * gw : 1;
  }
  long long cells() const { return (long long)is * js * ks; }
  long long at(int i, int j, int k, int v) const {
    return (((long long)v * ks + k) * js + j) * (long long)is + i;
  }
};

}  // namespace

// counter-based generator for the random KH layer: splitmix64 of (seed, global cell, draw).
// The reference draws from a Kokkos XorShift64 pool whose stream depends on the thread
// mapping, so no generator reproduces it; this one is identical on every device.
double kh_uniform(uint64_t seed, uint64_t cell, uint64_t draw) {
  uint64_t x = seed * 0x9E3779B97F4A7C15ull + cell * 0xD1B54A32D192ED03ull + draw * 0x8CB92BA72F3D8DD7ull;
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  x = x ^ (x >> 31);
  return (double)(x >> 11) * (1.0 / 9007199254740992.0);
}

void init_problem_host(const ppk_mhd_params &p, const ppk_mhd_problem &pr, double *U) {
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
  }
}


}  // namespace ppk
