// Initial conditions of the MHD problems, computed on the host in the reference layout
// (then uploaded), from the per-cell functions of ics_body.hpp that the on-device
// initialisation (ics_device.hip) shares.  Restates src/muscl/MHDInitFunctors2D.h and
// src/muscl/MHDInitFunctors3D.h (Orszag-Tang, blast, Kelvin-Helmholtz, linear wave,
// implode, rotor, field loop); citations per branch in ics_body.hpp.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "capi_internal.hpp"
#include "ics_body.hpp"
#include "ppkmhd_amd.h"

namespace ppk {

icb::IcConsts ic_consts(const ppk_mhd_params &p, const ppk_mhd_problem &pr) {
  icb::IcConsts K{};
  const double gamma0 = p.gamma0;
  const double PI = icb::kPi, TWOPI = 2 * PI;
  if (p.dim == 2) {  // InitOrszagTangFunctor2D, src/muscl/MHDInitFunctors2D.h:289-383
    K.B0 = 1.0 / std::sqrt(2 * TWOPI) * pr.ot_b0_scale;
    K.p0 = gamma0 / (2 * TWOPI);
  } else {  // InitOrszagTangFunctor3D, src/muscl/MHDInitFunctors3D.h:311-409
    K.B0 = 1.0 / std::sqrt(4 * PI) * pr.ot_b0_scale;
    K.p0 = gamma0 / (4 * PI);
  }
  K.d0 = gamma0 * K.p0;
  const double TwoPi = 4.0 * std::asin(1.0);
  K.p0e = gamma0 / (2.0 * TwoPi);
  if (pr.kind == PPK_PROBLEM_WAVE) {
    // WaveParams, src/shared/problems/WaveParams.h:31-146
    static const double revs[4][7] = {
        {4.472136e-01, -8.944272e-01, 4.216370e-01, 1.490712e-01, 2.012461e+00, 8.432740e-01,
         2.981424e-01},
        {0.0, 0.0, -3.333333e-01, 9.428090e-01, 0.0, -3.333333e-01, 9.428090e-01},
        {8.944272e-01, -4.472136e-01, -8.432740e-01, -2.981424e-01, 6.708204e-01,
         -4.216370e-01, -1.490712e-01},
        {1.0, 1.0, 0.0, 0.0, 0.5, 0.0, 0.0}};
    const int t = (pr.wave_type >= 0 && pr.wave_type <= 2) ? pr.wave_type : 3;
    std::memcpy(K.rev, revs[t], sizeof K.rev);
    K.V0 = t == 3 ? 1.0 : 0.0;
    K.amp = pr.wave_amplitude;
    K.wd0 = 1.0;
    K.wp0 = 1.0 / pr.wave_gamma0;
    const double Lx = pr.wave_lx, Ly = pr.wave_ly, Lz = pr.wave_lz;
    const double ang_3 = std::atan(Lx / Ly);
    K.sin_a3 = std::sin(ang_3);
    K.cos_a3 = std::cos(ang_3);
    const double ang_2 = std::atan(0.5 * (Lx * K.cos_a3 + Ly * K.sin_a3) / Lz);
    K.sin_a2 = std::sin(ang_2);
    K.cos_a2 = std::cos(ang_2);
    const double x1 = Lx * K.cos_a2 * K.cos_a3, x2 = Ly * K.cos_a2 * K.sin_a3, x3 = Lz * K.sin_a2;
    double lambda = x1;
    if (ang_3 != 0.0) lambda = std::fmin(lambda, x2);
    if (ang_2 != 0.0) lambda = std::fmin(lambda, x3);
    K.k_par = TwoPi / lambda;
    K.dby = K.amp * K.rev[5];
    K.dbz = K.amp * K.rev[6];
    K.bx0 = 1.0;
    K.by0 = std::sqrt(2.0);
    K.bz0 = 0.5;
  }
  return K;
}

void init_problem_host(const ppk_mhd_params &p, const ppk_mhd_problem &pr, double *U) {
  const icb::IcBox b = icb::box_of(p);
  const icb::IcConsts K = ic_consts(p, pr);
  std::fill(U, U + b.cells() * 8, 0.0);  // Kokkos views are zero-initialized
  if (pr.kind == PPK_PROBLEM_WAVE) {
    if (p.dim == 2) return;  // InitWaveFunctor2D_MHD does nothing (:972-975)
    std::vector<double> A((size_t)b.cells() * 3, 0.0);
    for (int k = 0; k < b.ks; ++k)
      for (int j = 0; j < b.js; ++j)
        for (int i = 0; i < b.is; ++i) icb::ic_wave_potential(p, K, b, i, j, k, A.data());
    for (int k = b.gw - 1; k < b.ks - b.gw + 1; ++k)
      for (int j = b.gw - 1; j < b.js - b.gw + 1; ++j)
        for (int i = b.gw - 1; i < b.is - b.gw + 1; ++i) icb::ic_wave_field(p, b, i, j, k, A.data(), U);
    for (int k = b.gw; k < b.ks - b.gw; ++k)
      for (int j = b.gw; j < b.js - b.gw; ++j)
        for (int i = b.gw; i < b.is - b.gw; ++i) icb::ic_wave_state(p, K, b, i, j, k, U);
    return;
  }
  if (pr.kind == PPK_PROBLEM_FIELD_LOOP) {
    std::vector<double> A((size_t)b.cells() * 3, 0.0);
    for (int k = 0; k < b.ks; ++k)
      for (int j = 0; j < b.js; ++j)
        for (int i = 0; i < b.is; ++i) icb::ic_loop_potential(p, pr, b, i, j, k, A.data());
    const int k0 = p.dim == 3 ? b.gw : 0, k1 = p.dim == 3 ? b.ks - b.gw : 1;
    for (int k = k0; k < k1; ++k)
      for (int j = b.gw; j < b.js - b.gw; ++j)
        for (int i = b.gw; i < b.is - b.gw; ++i) icb::ic_loop_state(p, pr, b, i, j, k, A.data(), U);
    for (int k = k0; k < k1; ++k)
      for (int j = b.gw; j < b.js - b.gw; ++j)
        for (int i = b.gw; i < b.is - b.gw; ++i) icb::ic_loop_energy(p, b, i, j, k, U);
    return;
  }
  for (int k = 0; k < b.ks; ++k)
    for (int j = 0; j < b.js; ++j)
      for (int i = 0; i < b.is; ++i) icb::ic_cell(p, pr, K, b, i, j, k, U);
  if (pr.kind == PPK_PROBLEM_ORSZAG_TANG)
    for (int k = 0; k < b.ks; ++k)
      for (int j = 0; j < b.js - 1; ++j)
        for (int i = 0; i < b.is - 1; ++i) icb::ic_ot_energy(p, K, b, i, j, k, U);
}

}  // namespace ppk
