// C ABI of the Euler (hydro) solver: ppk_hydro_* (include/ppkmhd_amd.h).  Same handle model
// and device-resident time loop as the MHD solver (ctrl block, per-block CFL partial maxima),
// with the hydro kernels of hydro.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <new>
#include <string>
#include <vector>

#include "capi_internal.hpp"
#include "event_timer.hpp"
#include "hydro_kernels.hpp"
#include "ics_body.hpp"
#include "ppkmhd_amd.h"

namespace ppk {
namespace {

// stages of the hydro step bracketed by the optional timers (ppk_hydro_timers_*)
enum HydroStage { HS_BOUNDARIES = 0, HS_PRIM_DT, HS_FLUXES, HS_UPDATE, HS_FUSED, HS_CONTROL,
                  HS_COUNT };
const char *const kHydroStageNames[HS_COUNT] = {"boundaries", "primitives_dt", "fluxes",
                                                "update", "fused", "control"};
static_assert(HS_COUNT <= ST_COUNT, "the event timer holds ST_COUNT stages");

const HydroTable kHydroFast = {fast::hydro_boundaries,    fast::hydro_prim_dt,
                               fast::hydro_prim_parts,    fast::hydro_staged_fluxes,
                               fast::hydro_staged_update, fast::hydro_fused,
                               fast::hydro_fused_parts};
const HydroTable kHydroStrict = {strict::hydro_boundaries,    strict::hydro_prim_dt,
                                 strict::hydro_prim_parts,    strict::hydro_staged_fluxes,
                                 strict::hydro_staged_update, strict::hydro_fused,
                                 strict::hydro_fused_parts};

int validate_hydro(const ppk_mhd_params *p, std::string &why) {
  if (!p) { why = "params is NULL"; return PPK_ERR_INVALID_ARGUMENT; }
  if (p->dim != 2 && p->dim != 3) { why = "dim must be 2 or 3"; return PPK_ERR_INVALID_ARGUMENT; }
  if (p->implementation_version < 0 || p->implementation_version > 2) {
    why = "implementationVersion must be 0, 1 or 2 (HydroParams.cpp:446-451)";
    return PPK_ERR_INVALID_ARGUMENT;
  }
  if (p->implementation_version == 2) {  // SolverHydroMuscl has versions 0 and 1 only
    why = "implementationVersion 2 has no hydro scheme (SolverHydroMuscl.cpp:277-319)";
    return PPK_ERR_UNSUPPORTED;
  }
  if (p->ghost_width != 2) {
    why = "ghost_width must be 2 for the hydro solvers (HydroParams.cpp:42-55)";
    return PPK_ERR_INVALID_ARGUMENT;
  }
  if (p->nx < 2 || p->ny < 2 || (p->dim == 3 && p->nz < 2)) {
    why = "nx, ny (and nz in 3D) must be >= ghost_width";
    return PPK_ERR_INVALID_ARGUMENT;
  }
  if (p->dim == 2 && p->nz != 1) { why = "nz must be 1 in 2D"; return PPK_ERR_INVALID_ARGUMENT; }
  const long long cells = (long long)(p->nx + 4) * (p->ny + 4) * (p->dim == 3 ? p->nz + 4 : 1);
  if (cells > (1LL << 31)) { why = "grid too large (more than 2^31 padded cells)"; return PPK_ERR_INVALID_ARGUMENT; }
  // riemann_hydro (RiemannSolvers.h:573-633): approx, hll (2D only), hllc, llf
  const bool ok_riemann = p->riemann == PPK_RIEMANN_APPROX || p->riemann == PPK_RIEMANN_LLF ||
                          p->riemann == PPK_RIEMANN_HLLC ||
                          (p->riemann == PPK_RIEMANN_HLL && p->dim == 2);
  if (!ok_riemann) {
    why = p->riemann == PPK_RIEMANN_HLL
              ? "riemann hll has no 3D hydro solver in the reference (RiemannSolvers.h:609-633)"
              : "riemann must be approx, llf, hllc (or hll in 2D) for the hydro solvers";
    return PPK_ERR_UNSUPPORTED;
  }
  if (p->slope_type != 0.0 && p->slope_type != 1.0 && p->slope_type != 2.0) {
    why = "slope_type must be 0, 1 or 2 (HydroBaseFunctor2D.h:470-510)";
    return PPK_ERR_UNSUPPORTED;
  }
  const int nf = p->dim == 3 ? 6 : 4;
  for (int f = 0; f < nf; ++f)
    if (p->bc[f] != PPK_BC_DIRICHLET && p->bc[f] != PPK_BC_NEUMANN && p->bc[f] != PPK_BC_PERIODIC) {
      why = "boundary types must be 1 (Dirichlet), 2 (Neumann) or 3 (periodic)";
      return PPK_ERR_INVALID_ARGUMENT;
    }
  if (p->precision != PPK_FP64 && p->precision != PPK_FP64_STRICT) {
    why = p->precision == PPK_FP32_BULK ? "the hydro solvers compute in FP64 only"
                                        : "unknown precision";
    return p->precision == PPK_FP32_BULK ? PPK_ERR_UNSUPPORTED : PPK_ERR_INVALID_ARGUMENT;
  }
  if (p->kernel_path < PPK_PATH_AUTO || p->kernel_path > PPK_PATH_FUSED) {
    why = "unknown kernel_path";
    return PPK_ERR_INVALID_ARGUMENT;
  }
  if (!(p->gamma0 > 1.0) || !(p->dx > 0) || !(p->dy > 0) || (p->dim == 3 && !(p->dz > 0))) {
    why = "gamma0 must be > 1 and dx, dy, dz > 0 (call ppk_mhd_params_init)";
    return PPK_ERR_INVALID_ARGUMENT;
  }
  return PPK_OK;
}

HydroDev make_hydro_dev(const ppk_mhd_params &p) {
  HydroDev d{};
  d.dim = p.dim;
  d.nv = p.dim == 3 ? 5 : 4;
  d.nx = p.nx;
  d.ny = p.ny;
  d.nz = p.dim == 3 ? p.nz : 1;
  d.gw = p.ghost_width;
  d.isize = p.nx + 2 * d.gw;
  d.jsize = p.ny + 2 * d.gw;
  d.ksize = p.dim == 3 ? p.nz + 2 * d.gw : 1;
  for (int f = 0; f < 6; ++f) d.bc[f] = p.bc[f];
  d.plane = (long long)d.isize * d.jsize;
  d.vstride = d.plane * d.ksize;
  d.dx = p.dx;
  d.dy = p.dy;
  d.dz = p.dz;
  d.idx = 1.0 / p.dx;
  d.idy = 1.0 / p.dy;
  d.idz = p.dim == 3 ? 1.0 / p.dz : 0.0;
  d.P.gamma0 = p.gamma0;
  d.P.gamma6 = p.gamma6;
  d.P.smallr = p.smallr;
  d.P.smallc = p.smallc;
  d.P.smallp = p.smallp;
  d.P.smallpp = p.smallpp;
  d.P.slope_type = p.slope_type;
  d.P.riemann = p.riemann;
  d.P.entho = 1.0 / (p.gamma0 - 1.0);  // RiemannSolvers.h: entho = ONE_F / (gamma0 - ONE_F)
  // gravity; the field (allocated by the handle) is set by the problem, zero until then as
  // the reference's
  d.grav = p.gravity_static ? 1 : 0;
  d.gf = nullptr;
  d.cfl = p.cfl;
  d.dxmin = p.dim == 3 ? std::fmin(std::fmin(p.dx, p.dy), p.dz) : std::fmin(p.dx, p.dy);
  d.idxmin = 1.0 / d.dxmin;
  // tuning knob: output rows (2D) / planes (3D) of one marching wave / block
  const char *strip = std::getenv("PPKMHD_HYDRO_STRIP");
  const int rows = strip ? std::atoi(strip) : 0;
  d.strip_rows = rows > 0 ? rows : (p.dim == 3 ? 48 : 64);  // measured best at 8192 x 4096 / 512 x 384 x 256
  return d;
}

// The 19 four-quadrant 2D Riemann configurations of Lax & Liu (SIAM J. Sci. Comput. 19(2),
// 1998), primitive (rho, u, v, p) per quadrant, quadrants numbered counter-clockwise from the
// upper right (x >= xc, y >= yc) as the reference's U0..U3
// (src/shared/problems/initRiemannConfig2d.h:42-514).  Values as the reference holds them,
// including configuration 6's quadrant 1 at p = 0.5.
constexpr double kLaxLiu[19][4][4] = {
    {{1.0, 0.0, 0.0, 1.0}, {0.5197, -0.7259, 0.0, 0.4}, {0.1072, -0.7259, -1.4045, 0.0439},
     {0.2579, 0.0, -1.4045, 0.15}},
    {{1.0, 0.0, 0.0, 1.0}, {0.5197, -0.7259, 0.0, 0.4}, {1.0, -0.7259, -0.7259, 1.0},
     {0.5197, 0.0, -0.7259, 0.4}},
    {{1.5, 0.0, 0.0, 1.5}, {0.5323, 1.206, 0.0, 0.3}, {0.138, 1.206, 1.206, 0.029},
     {0.5323, 0.0, 1.206, 0.3}},
    {{1.1, 0.0, 0.0, 1.1}, {0.5065, 0.8939, 0.0, 0.35}, {1.1, 0.8939, 0.8939, 1.1},
     {0.5065, 0.0, 0.8939, 0.35}},
    {{1.0, -0.75, -0.5, 1.0}, {2.0, -0.75, 0.5, 1.0}, {1.0, 0.75, 0.5, 1.0},
     {3.0, 0.75, -0.5, 1.0}},
    {{1.0, 0.75, -0.5, 1.0}, {2.0, 0.75, 0.5, 0.5}, {1.0, -0.75, 0.5, 1.0},
     {3.0, -0.75, -0.5, 1.0}},
    {{1.0, 0.1, 0.1, 1.0}, {0.5197, -0.6259, 0.1, 0.4}, {0.8, 0.1, 0.1, 0.4},
     {0.5197, 0.1, -0.6259, 0.4}},
    {{0.5197, 0.1, 0.1, 0.4}, {1.0, -0.6259, 0.1, 1.0}, {0.8, 0.1, 0.1, 1.0},
     {1.0, 0.1, -0.6259, 1.0}},
    {{1.0, 0.0, 0.3, 1.0}, {2.0, 0.0, -0.3, 1.0}, {1.039, 0.0, -0.8133, 0.4},
     {0.5197, 0.0, -0.4259, 0.4}},
    {{1.0, 0.0, 0.4297, 1.0}, {0.5, 0.0, 0.6076, 1.0}, {0.2281, 0.0, -0.6076, 0.3333},
     {0.4562, 0.0, -0.4259, 0.3333}},
    {{1.0, 0.1, 0.0, 1.0}, {0.5313, 0.8276, 0.0, 0.4}, {0.8, 0.1, 0.0, 0.4},
     {0.5313, 0.1, 0.7276, 0.4}},
    {{0.5313, 0.0, 0.0, 0.4}, {1.0, 0.7276, 0.0, 1.0}, {0.8, 0.0, 0.0, 1.0},
     {1.0, 0.0, 0.7276, 1.0}},
    {{1.0, 0.0, -0.3, 1.0}, {2.0, 0.0, 0.3, 1.0}, {1.0625, 0.0, 0.8145, 0.4},
     {0.5313, 0.0, 0.4276, 0.4}},
    {{2.0, 0.0, -0.5606, 8.0}, {1.0, 0.0, -1.2172, 8.0}, {0.4736, 0.0, 1.2172, 2.6667},
     {0.9474, 0.0, 1.1606, 2.6667}},
    {{1.0, 0.1, -0.3, 1.0}, {0.5197, -0.6259, -0.3, 0.4}, {0.8, 0.1, -0.3, 0.4},
     {0.5313, 0.1, 0.4276, 0.4}},
    {{0.5313, 0.1, 0.1, 0.4}, {1.0222, -0.6179, 0.1, 1.0}, {0.8, 0.1, 0.1, 1.0},
     {1.0, 0.1, 0.8276, 1.0}},
    {{1.0, 0.0, -0.4, 1.0}, {2.0, 0.0, -0.3, 1.0}, {1.0625, 0.0, 0.2145, 0.4},
     {0.5197, 0.0, -1.1259, 0.4}},
    {{1.0, 0.0, 1.0, 1.0}, {2.0, 0.0, -0.3, 1.0}, {1.0625, 0.0, 0.2145, 0.4},
     {0.5197, 0.0, 0.2741, 0.4}},
    {{1.0, 0.0, 0.3, 1.0}, {2.0, 0.0, -0.3, 1.0}, {1.0625, 0.0, 0.2145, 0.4},
     {0.5197, 0.0, -0.4259, 0.4}},
};

// [riemann2d] config_number -> row of kLaxLiu.  The reference's selection (a switch on the
// number) has no break after its last case, so number 18 falls into the default, which is
// configuration 1 like every number outside 0..17; configuration 19 is never selected.
int lax_liu_row(int config_number) {
  return config_number >= 0 && config_number <= 17 ? config_number : 0;
}

// initial conditions of the hydro problems in the reference layout (host)
// InitImplodeFunctor2D/3D (src/muscl/HydroInitFunctors2D.h:35-125, HydroInitFunctors3D.h:66-165),
// InitBlastFunctor2D/3D (HydroInitFunctors2D.h:129-212, HydroInitFunctors3D.h:170-262),
// InitKelvinHelmholtzFunctor2D/3D (HydroInitFunctors2D.h:216-348, 3D.h:271-427),
// InitGreshoVortexFunctor2D/3D (HydroInitFunctors2D.h:352-449, 3D.h:429-530),
// InitIsentropicVortexFunctor2D (HydroInitFunctors2D.h:570-660),
// InitFourQuadrantFunctor2D (HydroInitFunctors2D.h:453-565, SolverHydroMuscl.cpp:72-89),
// RayleighTaylorInstabilityFunctor2D/3D (HydroInitFunctors2D.h:663-787, 3D.h:548-677),
// RisingBubbleFunctor2D/3D (HydroInitFunctors2D.h:790-906, 3D.h:685-802)
void hydro_ics(const ppk_mhd_params &p, const ppk_mhd_problem &pr, double *U) {
  const int gw = p.ghost_width, dim = p.dim, nv = dim == 3 ? 5 : 4;
  const int is = p.nx + 2 * gw, js = p.ny + 2 * gw, ks = dim == 3 ? p.nz + 2 * gw : 1;
  const long long vs = (long long)is * js * ks;
  std::fill(U, U + nv * vs, 0.0);
  const double g0 = p.gamma0;
  for (int k = 0; k < ks; ++k)
    for (int j = 0; j < js; ++j)
      for (int i = 0; i < is; ++i) {
        const long long c = ((long long)k * js + j) * is + i;
        const double x = p.xmin + p.dx / 2 + (i - gw) * p.dx;
        const double y = p.ymin + p.dy / 2 + (j - gw) * p.dy;
        const double z = dim == 3 ? p.zmin + p.dz / 2 + (k - gw) * p.dz : 0.0;
        double s[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        if (pr.kind == PPK_PROBLEM_KELVIN_HELMHOLTZ) {
          // the random layer uses the library's counter generator (splitmix64 of seed, cell,
          // draw), as the MHD Kelvin-Helmholtz problem does; the reference's XorShift pool
          // depends on the thread mapping
          const double d_in = pr.kh_d_in, d_out = pr.kh_d_out, vin = pr.kh_vflow_in;
          const double vout = pr.kh_vflow_out, ampl = pr.kh_amplitude, pres = pr.kh_pressure;
          const uint64_t cell = (uint64_t)c;
          if (pr.kh_p_rand) {
            const double sn = dim == 2 ? (y - p.ymin) / (p.ymax - p.ymin)
                                       : (z - p.zmin) / (p.zmax - p.zmin);
            const bool outer = sn < 0.25 || sn > 0.75;
            const double d = outer ? d_out : d_in;
            double u = outer ? vout : vin, v = 0.0, w = 0.0;
            u += ampl * (icb::kh_uniform(pr.kh_seed, cell, 0) - 0.5);
            v += ampl * (icb::kh_uniform(pr.kh_seed, cell, 1) - 0.5);
            if (dim == 3) w += ampl * (icb::kh_uniform(pr.kh_seed, cell, 2) - 0.5);
            s[ID] = d;
            s[IU] = d * u;
            s[IV] = d * v;
            s[IW] = d * w;
            s[IP] = dim == 2 ? pres / (g0 - 1.0) + 0.5 * d * (u * u + v * v)
                             : pres / (g0 - 1.0) + 0.5 * d * (u * u + v * v + w * w);
          } else if (pr.kh_p_sine_rob) {
            const int n = pr.kh_mode;
            const double w0 = pr.kh_w0, delta = pr.kh_delta, s1 = 0.25, s2 = 0.75;
            const double sc = dim == 2 ? y : z;
            const double ramp = 1.0 / (1.0 + std::exp(2 * (sc - s1) / delta)) +
                                1.0 / (1.0 + std::exp(2 * (s2 - sc) / delta));
            const double d = d_in + ramp * (d_out - d_in);
            const double u = vin + ramp * (vout - vin);
            if (dim == 2) {
              const double v = w0 * std::sin(n * icb::kPi * x);
              s[ID] = d;
              s[IU] = d * u;
              s[IV] = d * v;
              s[IP] = pres / (g0 - 1.0) + 0.5 * d * (u * u + v * v);
            } else {
              const double v1y = vin / 2, v2y = vout / 2;
              const double v = v1y + ramp * (v2y - v1y);
              const double w = w0 * std::sin(n * icb::kPi * x) * std::sin(n * icb::kPi * y);
              s[ID] = d;
              s[IU] = d * u;
              s[IV] = d * v;
              s[IW] = d * w;
              s[IP] = pres / (g0 - 1.0) + 0.5 * d * (u * u + v * v + w * w);
            }
          }  // neither perturbation: the reference leaves the cell at zero
        } else if (pr.kind == PPK_PROBLEM_GRESHO_VORTEX) {
          const double rho0 = pr.gresho_rho0, Ma = pr.gresho_ma;
          const double p0 = rho0 / (g0 * Ma * Ma);
          const double r = std::sqrt(x * x + y * y);
          const double theta = std::atan2(y, x);
          const double cosT = std::cos(theta), sinT = std::sin(theta);
          double uphi, pp;
          if (r < 0.2) {
            uphi = 5 * r;
            pp = p0 + 25 / 2.0 * r * r;
          } else if (r < 0.4) {
            uphi = 2 - 5 * r;
            pp = p0 + 25 / 2.0 * r * r + 4 * (1 - 5 * r - std::log(0.2) + std::log(r));
          } else {
            uphi = 0;
            pp = p0 - 2 + 4 * std::log(2.0);
          }
          s[ID] = rho0;
          s[IU] = rho0 * (-sinT * uphi);
          s[IV] = rho0 * (cosT * uphi);
          if (dim == 2) {
            s[IP] = pp / (g0 - 1.0) + 0.5 * (s[IU] * s[IU] + s[IV] * s[IV]) / s[ID];
          } else {
            s[IW] = rho0 * (cosT * uphi);  // as the reference's 3D functor writes it
            s[IP] = pp / (g0 - 1.0) + 0.5 * (s[IU] * s[IU] + s[IV] * s[IV] + s[IW] * s[IW]) / s[ID];
          }
        } else if (pr.kind == PPK_PROBLEM_ISENTROPIC_VORTEX) {  // 2D
          const double xp = x - pr.iv_x, yp = y - pr.iv_y;
          const double r = std::sqrt(xp * xp + yp * yp);
          const double beta = pr.iv_beta, T_a = pr.iv_t_a;
          const double du = -yp * beta / (2 * icb::kPi) * std::exp(0.5 * (1.0 - r * r));
          const double dv = xp * beta / (2 * icb::kPi) * std::exp(0.5 * (1.0 - r * r));
          const double T = T_a - (g0 - 1) * beta * beta / (8 * g0 * icb::kPi * icb::kPi) *
                                     std::exp(1.0 - r * r);
          const double rho = pr.iv_rho_a * std::pow(T / T_a, 1.0 / (g0 - 1));
          s[ID] = rho;
          s[IU] = rho * (pr.iv_u_a + du);
          s[IV] = rho * (pr.iv_v_a + dv);
          s[IP] = rho * T / (g0 - 1.0) + 0.5 * rho * (pr.iv_u_a + du) * (pr.iv_u_a + du) +
                  0.5 * rho * (pr.iv_v_a + dv) * (pr.iv_v_a + dv);
        } else if (pr.kind == PPK_PROBLEM_RAYLEIGH_TAYLOR) {
          // heavy fluid over light: density d1 above the mid-plane of the vertical axis (y in
          // 2D, z in 3D), a cosine bump of the vertical velocity written into the momentum
          // slot as the reference does, pressure P0 + rho (g . x) with P0 = 1 (hydrostatic)
          const double P0 = 1.0, Lx = p.xmax - p.xmin, Ly = p.ymax - p.ymin;
          const double bump = pr.rti_amplitude * (1 + std::cos(2 * icb::kPi * x / Lx)) *
                              (1 + std::cos(2 * icb::kPi * y / Ly)) / 4;
          if (dim == 2) {
            s[ID] = y > (p.ymin + p.ymax) / 2 ? pr.rti_d1 : pr.rti_d0;
            s[IU] = 0.0;
            s[IV] = bump;
            s[IP] = (P0 + s[ID] * (pr.rti_gx * x + pr.rti_gy * y)) / (g0 - 1.0);
          } else {
            s[ID] = z > (p.zmin + p.zmax) / 2 ? pr.rti_d1 : pr.rti_d0;
            s[IU] = 0.0;
            s[IV] = 0.0;
            s[IW] = bump;
            s[IP] = (P0 + s[ID] * (pr.rti_gx * x + pr.rti_gy * y + pr.rti_gz * z)) / (g0 - 1);
          }
        } else if (pr.kind == PPK_PROBLEM_RISING_BUBBLE) {
          // a ball of density din in dout, at rest, pressure P0 + rho (g . x) with P0 = 1
          const double P0 = 1.0;
          double r2 = (x - pr.rb_x0) * (x - pr.rb_x0) + (y - pr.rb_y0) * (y - pr.rb_y0);
          if (dim == 3) r2 = r2 + (z - pr.rb_z0) * (z - pr.rb_z0);
          s[ID] = std::sqrt(r2) < pr.rb_R ? pr.rb_din : pr.rb_dout;
          s[IU] = s[IV] = s[IW] = 0.0;
          s[IP] = dim == 2 ? (P0 + s[ID] * (pr.rb_gx * x + pr.rb_gy * y)) / (g0 - 1.0)
                           : (P0 + s[ID] * (pr.rb_gx * x + pr.rb_gy * y + pr.rb_gz * z)) /
                                 (g0 - 1.0);
        } else if (pr.kind == PPK_PROBLEM_DISK) {
          // a rotating disk around a softened point mass GM, in coordinates relative to the
          // disk centre (xc, yc, zc): power-law density with a vertical exponential profile,
          // divided by the contrast outside radius (with a ramp over contrast_width), sound
          // speed c(r) = c_ref / sqrt(r / radius) (r_inner inside), rotation from the radial
          // balance, pressure c^2 rho; the energy adds 0.5 rho omega2 (r^2 + eps^2)
          const double xr = x - pr.disk_xc, yr = y - pr.disk_yc;
          const double zr = dim == 3 ? z - pr.disk_zc : 0.0;
          const double GM = pr.pg_gm, eps = pr.pg_eps, R = pr.disk_radius;
          const double Ri = pr.disk_radius_inner, cw = pr.disk_contrast_width;
          const double r2 = xr * xr + yr * yr, r = std::sqrt(r2);
          const double r_soft = std::sqrt(r2 + eps * eps);
          const double r_sph = std::sqrt(r2 + eps * eps + zr * zr);
          const double x_soft = xr * (r_soft / r), y_soft = yr * (r_soft / r);
          const double cs = r < Ri ? pr.disk_ref_sound_speed / std::sqrt(Ri / R)
                                   : pr.disk_ref_sound_speed / std::sqrt(r / R);
          const double omega2 =
              r >= Ri ? GM / (r_soft * r_soft * r_sph) - 3.5 * (cs / r_soft) * (cs / r_soft)
                      : GM / (r_soft * r_soft * r_soft) - 2.5 * (cs / r_soft) * (cs / r_soft);
          const double omega = std::sqrt(std::fmax(omega2, 0.0));
          double rho = pr.disk_ref_density * std::pow(r_soft / R, -2.5) *
                       std::exp(GM / (cs * cs) * (1. / r_sph - 1. / r_soft));
          if (r > R || std::fabs(zr) > 0.5 * R) {
            if (r > R + cw || std::fabs(zr) > 0.5 * R + cw)
              rho /= pr.disk_contrast_density;
            else
              rho *= std::pow(pr.disk_contrast_density,
                              -std::fmax((r - R) / cw, (std::fabs(zr) - 0.5 * R) / cw));
          }
          rho = std::fmax(rho, 1e-6);
          s[ID] = rho;
          s[IU] = -y_soft * omega * rho;
          s[IV] = x_soft * omega * rho;
          s[IW] = 0.0;
          s[IP] = cs * cs * rho / (g0 - 1.0) + 0.5 * rho * omega2 * (r2 + eps * eps);
        } else if (pr.kind == PPK_PROBLEM_FOUR_QUADRANT) {  // 2D
          const int quad = x < pr.fq_x ? (y < pr.fq_y ? 2 : 1) : (y < pr.fq_y ? 3 : 0);
          const double *q = kLaxLiu[lax_liu_row(pr.fq_config)][quad];  // rho, u, v, p
          s[ID] = q[0];
          s[IU] = q[1] * q[0];
          s[IV] = q[2] * q[0];
          s[IP] = q[3] / (g0 - 1.0) + q[0] * (q[1] * q[1] + q[2] * q[2]) * 0.5;
        } else if (pr.kind == PPK_PROBLEM_BLAST) {
          double d2 = (x - pr.blast_center_x) * (x - pr.blast_center_x) +
                      (y - pr.blast_center_y) * (y - pr.blast_center_y);
          if (dim == 3) d2 = d2 + (z - pr.blast_center_z) * (z - pr.blast_center_z);
          const bool in = d2 < pr.blast_radius * pr.blast_radius;
          s[ID] = in ? pr.blast_density_in : pr.blast_density_out;
          s[IP] = (in ? pr.blast_pressure_in : pr.blast_pressure_out) / (g0 - 1.0);
          s[IU] = s[IV] = s[IW] = 0.0;
        } else {
          bool outer;
          if (dim == 2)
            outer = pr.implode_shape == 1 ? (x + y * y > 0.5 && x + y * y < 1.5)
                                          : (x + y > (p.xmin + p.xmax) / 2. + p.ymin);
          else
            outer = pr.implode_shape == 1
                        ? (x + y + z > 0.5 && x + y + z < 2.5)
                        : (x + y + z > (p.xmin + p.xmax) / 2. + p.ymin + p.zmin);
          const double *st = outer ? pr.implode_out : pr.implode_in;  // rho, p, u, v, w
          const double rho = st[0], pre = st[1], u = st[2], v = st[3], w = st[4];
          s[ID] = rho;
          s[IP] = dim == 2 ? pre / (g0 - 1.0) + 0.5 * rho * (u * u + v * v)
                           : pre / (g0 - 1.0) + 0.5 * rho * (u * u + v * v + w * w);
          // the reference stores the velocities in the momentum slots
          s[IU] = u;
          s[IV] = v;
          s[IW] = w;
        }
        for (int n = 0; n < nv; ++n) U[c + n * vs] = s[n];
      }
}

// Gravity field of a problem, dim planes of the padded box: the uniform field of
// rayleigh_taylor / rising_bubble, the softened point mass of disk
// (PointSourceGravity::eval, src/shared/problems/PointSourceGravity.h, evaluated by the disk
// functor at coordinates relative to the disk centre), zero otherwise.
void hydro_gravity_field(const ppk_mhd_params &p, const ppk_mhd_problem &pr, double *G) {
  const int gw = p.ghost_width, dim = p.dim;
  const int is = p.nx + 2 * gw, js = p.ny + 2 * gw, ks = dim == 3 ? p.nz + 2 * gw : 1;
  const long long vs = (long long)is * js * ks;
  for (int k = 0; k < ks; ++k)
    for (int j = 0; j < js; ++j)
      for (int i = 0; i < is; ++i) {
        const long long c = ((long long)k * js + j) * is + i;
        double g[3] = {0.0, 0.0, 0.0};
        if (pr.kind == PPK_PROBLEM_RAYLEIGH_TAYLOR) {
          g[0] = pr.rti_gx; g[1] = pr.rti_gy; g[2] = pr.rti_gz;
        } else if (pr.kind == PPK_PROBLEM_RISING_BUBBLE) {
          g[0] = pr.rb_gx; g[1] = pr.rb_gy; g[2] = pr.rb_gz;
        } else if (pr.kind == PPK_PROBLEM_DISK) {
          const double x = p.xmin + p.dx / 2 + (i - gw) * p.dx - pr.disk_xc;
          const double y = p.ymin + p.dy / 2 + (j - gw) * p.dy - pr.disk_yc;
          const double z = dim == 3 ? p.zmin + p.dz / 2 + (k - gw) * p.dz - pr.disk_zc : 0.0;
          const double r = std::sqrt((x - pr.pg_x) * (x - pr.pg_x) + (y - pr.pg_y) * (y - pr.pg_y) +
                                     (z - pr.pg_z) * (z - pr.pg_z) + pr.pg_eps * pr.pg_eps);
          const double GM2 = pr.pg_gm / (r * r * r);
          g[0] = GM2 * (pr.pg_x - x);
          g[1] = GM2 * (pr.pg_y - y);
          g[2] = GM2 * (pr.pg_z - z);
        }
        for (int d = 0; d < dim; ++d) G[c + d * vs] = g[d];
      }
}

}  // namespace
}  // namespace ppk

struct ppk_hydro_solver {
  ppk_mhd_params p{};
  ppk::HydroDev hd{};
  hipStream_t stream = nullptr;
  double *U[2] = {nullptr, nullptr};
  double *Q = nullptr;  // staged form: primitives
  double *F = nullptr;  // staged form: dim x nv flux planes
  double *G = nullptr;  // gravity: dim planes of the per-cell field
  double *ctrl = nullptr;
  double *ctrl_host = nullptr;
  double *parts = nullptr;
  ppk::EventTimer *timer = nullptr;
  bool fused = false;
  bool cfl_valid = false;  // parts holds the fused kernel's CFL maxima of cur()
  const ppk::KernelTable *K = nullptr;
  const ppk::HydroTable *H = nullptr;
  long long iter_host = 0;
  std::string err;

  ppk::StageTimer *tm() { return (timer && timer->mask()) ? timer : nullptr; }
  double *cur() { return U[iter_host & 1]; }
  double *nxt() { return U[(iter_host + 1) & 1]; }
  long long cells() const { return (long long)hd.isize * hd.jsize * hd.ksize; }
  size_t values() const { return (size_t)cells() * hd.nv; }
  size_t state_bytes() const { return values() * sizeof(double); }
};

namespace {

int hfail(ppk_hydro_t *h, int code, const std::string &msg) {
  if (h)
    h->err = msg;
  else
    ppk::set_thread_error(msg);
  return code;
}

#define HHIP(h, expr)                                                                     \
  do {                                                                                    \
    hipError_t _e = (expr);                                                               \
    if (_e != hipSuccess)                                                                 \
      return hfail((h), PPK_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e));  \
  } while (0)

int hwrite_ctrl(ppk_hydro_t *h) {
  HHIP(h, hipMemcpyAsync(h->ctrl, h->ctrl_host, sizeof(double) * ppk::CTRL_SLOTS,
                         hipMemcpyHostToDevice, h->stream));
  HHIP(h, hipStreamSynchronize(h->stream));
  return PPK_OK;
}

int hread_ctrl(ppk_hydro_t *h) {
  HHIP(h, hipMemcpyAsync(h->ctrl_host, h->ctrl, sizeof(double) * ppk::CTRL_SLOTS,
                         hipMemcpyDeviceToHost, h->stream));
  HHIP(h, hipStreamSynchronize(h->stream));
  return PPK_OK;
}

void hfree_all(ppk_hydro_t *h) {
  for (void *ptr : {(void *)h->U[0], (void *)h->U[1], (void *)h->Q, (void *)h->F,
                    (void *)h->G, (void *)h->ctrl, (void *)h->parts})
    if (ptr) (void)hipFree(ptr);
  delete h->timer;
  h->timer = nullptr;
  if (h->ctrl_host) (void)hipHostFree(h->ctrl_host);
  if (h->stream) (void)hipStreamDestroy(h->stream);
}

// SolverBase::next_iteration -> SolverHydroMuscl::godunov_unsplit_impl, one step cur -> nxt
int hlaunch_step(ppk_hydro_t *h, double dt, int dt_mode, int check_finished) {
  const ppk::KernelTable &K = *h->K;
  const ppk::HydroTable &H = *h->H;
  const ppk::HydroDev &d = h->hd;
  hipStream_t s = h->stream;
  double *in = h->cur(), *out = h->nxt();
  const bool cfl = dt_mode == PPK_DT_CFL;
  const int mode = cfl ? ppk::SETDT_CFL : ppk::SETDT_FIXED;
  ppk::StageTimer *tm = h->tm();
  HHIP(h, K.begin(h->ctrl, check_finished, s));
  if (h->fused) {
    PPK_TIMED(tm, ppk::HS_BOUNDARIES, HHIP(h, H.boundaries(d, in, out, h->ctrl, s)));
    int nparts = H.fused_parts(d);
    if (cfl && !h->cfl_valid) {
      PPK_TIMED(tm, ppk::HS_PRIM_DT, HHIP(h, H.prim_dt(d, in, nullptr, h->parts, h->ctrl, 1, s)));
      nparts = H.prim_parts(d);
    }
    PPK_TIMED(tm, ppk::HS_CONTROL,
              HHIP(h, K.set_dt(h->ctrl, h->parts, nparts, dt, mode, h->p.cfl, s)));
    PPK_TIMED(tm, ppk::HS_FUSED, HHIP(h, H.fused(d, in, out, h->parts, h->ctrl, cfl ? 1 : 0, s)));
    h->cfl_valid = cfl;
  } else {
    PPK_TIMED(tm, ppk::HS_BOUNDARIES, HHIP(h, H.boundaries(d, in, nullptr, h->ctrl, s)));
    PPK_TIMED(tm, ppk::HS_PRIM_DT,
              HHIP(h, H.prim_dt(d, in, h->Q, h->parts, h->ctrl, cfl ? 1 : 0, s)));
    PPK_TIMED(tm, ppk::HS_CONTROL,
              HHIP(h, K.set_dt(h->ctrl, h->parts, H.prim_parts(d), dt, mode, h->p.cfl, s)));
    HHIP(h, hipMemcpyAsync(out, in, h->state_bytes(), hipMemcpyDeviceToDevice, s));
    PPK_TIMED(tm, ppk::HS_FLUXES, HHIP(h, H.staged_fluxes(d, h->Q, h->F, h->ctrl, s)));
    PPK_TIMED(tm, ppk::HS_UPDATE, HHIP(h, H.staged_update(d, h->F, out, h->ctrl, s)));
  }
  HHIP(h, K.end(h->ctrl, s));
  return PPK_OK;
}

// the rest of the constructor after init(U): make_boundaries(U), U2 = U, compute_dt
// (src/muscl/SolverHydroMuscl.h:289-300)
int hfinish_init(ppk_hydro_t *h) {
  h->ctrl_host[ppk::CTRL_T] = 0.0;
  h->ctrl_host[ppk::CTRL_ITER] = 0.0;
  h->ctrl_host[ppk::CTRL_ACTIVE] = 1.0;
  int rc = hwrite_ctrl(h);
  if (rc != PPK_OK) return rc;
  rc = ppk_hydro_make_boundaries(h);
  if (rc != PPK_OK) return rc;
  double dt = 0.0;
  rc = ppk_hydro_compute_dt(h, &dt);
  if (rc != PPK_OK) return rc;
  HHIP(h, hipMemcpyAsync(h->nxt(), h->cur(), h->state_bytes(), hipMemcpyDeviceToDevice,
                         h->stream));
  h->ctrl_host[ppk::CTRL_DT] = (0.0 + dt > h->p.tEnd) ? h->p.tEnd - 0.0 : dt;
  return hwrite_ctrl(h);
}

bool hydro_problem_ok(int kind, int dim) {
  return kind == PPK_PROBLEM_IMPLODE || kind == PPK_PROBLEM_BLAST ||
         kind == PPK_PROBLEM_KELVIN_HELMHOLTZ || kind == PPK_PROBLEM_GRESHO_VORTEX ||
         kind == PPK_PROBLEM_RAYLEIGH_TAYLOR || kind == PPK_PROBLEM_RISING_BUBBLE ||
         kind == PPK_PROBLEM_DISK ||
         ((kind == PPK_PROBLEM_ISENTROPIC_VORTEX || kind == PPK_PROBLEM_FOUR_QUADRANT) &&
          dim == 2);
}
const char *const kHydroProblems =
    "hydro problems: implode, blast, kelvin_helmholtz, gresho_vortex, isentropic_vortex (2D), "
    "four_quadrant (2D), rayleigh_taylor, rising_bubble, disk";
// the reference writes the gravity field of these problems into a view that exists only when
// gravity is enabled
bool hydro_problem_needs_gravity(int kind) {
  return kind == PPK_PROBLEM_RAYLEIGH_TAYLOR || kind == PPK_PROBLEM_RISING_BUBBLE ||
         kind == PPK_PROBLEM_DISK;
}
const char *const kNeedsGravity =
    "rayleigh_taylor, rising_bubble and disk need gravity ([gravity] static=yes or point=yes)";

}  // namespace

extern "C" {

const char *ppk_hydro_last_error(const ppk_hydro_t *h) {
  return h ? h->err.c_str() : ppk::thread_error();
}

int ppk_hydro_create(const ppk_mhd_params *params, ppk_hydro_t **handle) {
  if (!handle) return hfail(nullptr, PPK_ERR_INVALID_ARGUMENT, "handle is NULL");
  *handle = nullptr;
  std::string why;
  int rc = ppk::validate_hydro(params, why);
  if (rc != PPK_OK) return hfail(nullptr, rc, why);
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return hfail(nullptr, PPK_ERR_HIP, "no HIP device available");
  ppk_hydro_t *h = new (std::nothrow) ppk_hydro_t();
  if (!h) return hfail(nullptr, PPK_ERR_OUT_OF_MEMORY, "host allocation failed");
  h->p = *params;
  h->hd = ppk::make_hydro_dev(*params);
  const bool strict = params->precision == PPK_FP64_STRICT;
  h->K = strict ? &ppk::kStrict : &ppk::kFast;
  h->H = strict ? &ppk::kHydroStrict : &ppk::kHydroFast;
  h->fused = params->kernel_path != PPK_PATH_STAGED;
  auto bail = [&](int code, const std::string &m) {
    hfree_all(h);
    delete h;
    return hfail(nullptr, code, m);
  };
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess)
    return bail(PPK_ERR_HIP, "hipStreamCreateWithFlags failed");
  h->timer = new (std::nothrow) ppk::EventTimer(h->stream);
  if (!h->timer) return bail(PPK_ERR_OUT_OF_MEMORY, "host allocation failed");
  const size_t nvals = h->values();
  auto dalloc = [&](double **ptr, size_t count) {
    return hipMalloc((void **)ptr, count * sizeof(double)) == hipSuccess;
  };
  const size_t nparts = (size_t)std::max(h->H->prim_parts(h->hd), h->H->fused_parts(h->hd));
  bool ok = dalloc(&h->U[0], nvals) && dalloc(&h->U[1], nvals) &&
            dalloc(&h->ctrl, ppk::CTRL_SLOTS) && dalloc(&h->parts, nparts);
  if (ok && !h->fused) ok = dalloc(&h->Q, nvals) && dalloc(&h->F, (size_t)params->dim * nvals);
  if (ok && params->gravity_static) ok = dalloc(&h->G, (size_t)params->dim * h->cells());
  if (!ok) return bail(PPK_ERR_OUT_OF_MEMORY, "device allocation failed");
  h->hd.gf = h->G;
  if (h->G && hipMemsetAsync(h->G, 0, (size_t)params->dim * h->cells() * sizeof(double),
                             h->stream) != hipSuccess)
    return bail(PPK_ERR_HIP, "hipMemsetAsync failed");
  if (hipHostMalloc((void **)&h->ctrl_host, sizeof(double) * ppk::CTRL_SLOTS) != hipSuccess)
    return bail(PPK_ERR_OUT_OF_MEMORY, "pinned host allocation failed");
  if (hipMemsetAsync(h->U[0], 0, nvals * sizeof(double), h->stream) != hipSuccess ||
      hipMemsetAsync(h->U[1], 0, nvals * sizeof(double), h->stream) != hipSuccess)
    return bail(PPK_ERR_HIP, "hipMemsetAsync failed");
  std::fill(h->ctrl_host, h->ctrl_host + ppk::CTRL_SLOTS, 0.0);
  h->ctrl_host[ppk::CTRL_DT] = params->tEnd;
  h->ctrl_host[ppk::CTRL_TEND] = params->tEnd;
  h->ctrl_host[ppk::CTRL_NSTEPMAX] = (double)params->n_step_max;
  h->ctrl_host[ppk::CTRL_ACTIVE] = 1.0;
  if (hwrite_ctrl(h) != PPK_OK) {
    std::string m = h->err;
    return bail(PPK_ERR_HIP, m);
  }
  *handle = h;
  return PPK_OK;
}

void ppk_hydro_destroy(ppk_hydro_t *h) {
  if (!h) return;
  (void)hipStreamSynchronize(h->stream);
  hfree_all(h);
  delete h;
}

int ppk_hydro_nbvar(const ppk_hydro_t *h) { return h ? h->hd.nv : 0; }

int ppk_hydro_initial_state(const ppk_mhd_params *params, const ppk_mhd_problem *problem,
                            double *U, size_t count) {
  std::string why;
  int rc = ppk::validate_hydro(params, why);
  if (rc != PPK_OK) return hfail(nullptr, rc, why);
  if (!problem || !U) return hfail(nullptr, PPK_ERR_INVALID_ARGUMENT, "NULL argument");
  if (!hydro_problem_ok(problem->kind, params->dim))
    return hfail(nullptr, PPK_ERR_UNSUPPORTED, kHydroProblems);
  const int gw = params->ghost_width;
  const size_t cells = (size_t)(params->nx + 2 * gw) * (params->ny + 2 * gw) *
                       (params->dim == 3 ? params->nz + 2 * gw : 1);
  if (count != (size_t)(params->dim == 3 ? 5 : 4) * cells)
    return hfail(nullptr, PPK_ERR_INVALID_ARGUMENT, "count must be nbvar*isize*jsize*ksize");
  ppk::hydro_ics(*params, *problem, U);
  return PPK_OK;
}

int ppk_hydro_init_problem(ppk_hydro_t *h, const ppk_mhd_problem *problem) {
  if (!h || !problem) return hfail(h, PPK_ERR_INVALID_ARGUMENT, "NULL argument");
  if (!hydro_problem_ok(problem->kind, h->p.dim))
    return hfail(h, PPK_ERR_UNSUPPORTED, kHydroProblems);
  if (hydro_problem_needs_gravity(problem->kind) && !h->p.gravity_static)
    return hfail(h, PPK_ERR_INVALID_ARGUMENT, kNeedsGravity);
  std::vector<double> host(h->values());
  ppk::hydro_ics(h->p, *problem, host.data());
  if (h->G) {  // the problem's gravity field (zero for the problems that set none)
    std::vector<double> field((size_t)h->p.dim * h->cells());
    ppk::hydro_gravity_field(h->p, *problem, field.data());
    HHIP(h, hipMemcpyAsync(h->G, field.data(), field.size() * sizeof(double),
                           hipMemcpyHostToDevice, h->stream));
    HHIP(h, hipStreamSynchronize(h->stream));
  }
  h->iter_host = 0;
  h->cfl_valid = false;
  HHIP(h, hipMemcpyAsync(h->cur(), host.data(), h->state_bytes(), hipMemcpyHostToDevice,
                         h->stream));
  HHIP(h, hipStreamSynchronize(h->stream));
  return hfinish_init(h);
}

int ppk_hydro_upload(ppk_hydro_t *h, const double *U, size_t count) {
  if (!h || !U) return hfail(h, PPK_ERR_INVALID_ARGUMENT, "NULL argument");
  if (count != h->values())
    return hfail(h, PPK_ERR_INVALID_ARGUMENT, "count must be nbvar*isize*jsize*ksize");
  h->cfl_valid = false;
  HHIP(h, hipMemcpyAsync(h->cur(), U, h->state_bytes(), hipMemcpyHostToDevice, h->stream));
  HHIP(h, hipStreamSynchronize(h->stream));
  return PPK_OK;
}

int ppk_hydro_download(ppk_hydro_t *h, double *U, size_t count) {
  if (!h || !U) return hfail(h, PPK_ERR_INVALID_ARGUMENT, "NULL argument");
  if (count != h->values())
    return hfail(h, PPK_ERR_INVALID_ARGUMENT, "count must be nbvar*isize*jsize*ksize");
  HHIP(h, hipMemcpyAsync(U, h->cur(), h->state_bytes(), hipMemcpyDeviceToHost, h->stream));
  HHIP(h, hipStreamSynchronize(h->stream));
  return PPK_OK;
}

int ppk_hydro_device_state(ppk_hydro_t *h, double **dU) {
  if (!h || !dU) return hfail(h, PPK_ERR_INVALID_ARGUMENT, "NULL argument");
  h->cfl_valid = false;  // the caller may write the state
  *dU = h->cur();
  return PPK_OK;
}

int ppk_hydro_make_boundaries(ppk_hydro_t *h) {
  if (!h) return hfail(h, PPK_ERR_INVALID_ARGUMENT, "NULL handle");
  HHIP(h, h->K->begin(h->ctrl, 0, h->stream));
  HHIP(h, h->H->boundaries(h->hd, h->cur(), nullptr, h->ctrl, h->stream));
  return PPK_OK;
}

int ppk_hydro_compute_dt(ppk_hydro_t *h, double *dt) {
  if (!h || !dt) return hfail(h, PPK_ERR_INVALID_ARGUMENT, "NULL argument");
  // SolverHydroMuscl::compute_dt_local (SolverHydroMuscl.h:650-691): interior cells only
  HHIP(h, h->K->begin(h->ctrl, 0, h->stream));
  HHIP(h, h->H->prim_dt(h->hd, h->cur(), nullptr, h->parts, h->ctrl, 1, h->stream));
  HHIP(h, h->K->set_dt(h->ctrl, h->parts, h->H->prim_parts(h->hd), 0.0,
                       ppk::SETDT_REDUCE_ONLY, h->p.cfl, h->stream));
  h->cfl_valid = false;
  int rc = hread_ctrl(h);
  if (rc != PPK_OK) return rc;
  *dt = h->p.cfl / h->ctrl_host[ppk::CTRL_INVDT];
  return PPK_OK;
}

int ppk_hydro_step(ppk_hydro_t *h, double dt, int dt_mode) {
  if (!h) return hfail(h, PPK_ERR_INVALID_ARGUMENT, "NULL handle");
  if (dt_mode != PPK_DT_FIXED && dt_mode != PPK_DT_CFL)
    return hfail(h, PPK_ERR_INVALID_ARGUMENT, "dt_mode must be PPK_DT_FIXED or PPK_DT_CFL");
  int rc = hlaunch_step(h, dt, dt_mode, 0);
  if (rc != PPK_OK) return rc;
  h->iter_host += 1;
  return PPK_OK;
}

int ppk_hydro_run(ppk_hydro_t *h, int nsteps, double dt, int dt_mode) {
  if (!h) return hfail(h, PPK_ERR_INVALID_ARGUMENT, "NULL handle");
  if (nsteps < 0) return hfail(h, PPK_ERR_INVALID_ARGUMENT, "nsteps must be >= 0");
  if (dt_mode != PPK_DT_FIXED && dt_mode != PPK_DT_CFL)
    return hfail(h, PPK_ERR_INVALID_ARGUMENT, "dt_mode must be PPK_DT_FIXED or PPK_DT_CFL");
  const long long start = h->iter_host;
  const bool cfl_valid0 = h->cfl_valid;
  for (int s = 0; s < nsteps; ++s) {
    int rc = hlaunch_step(h, dt, dt_mode, 1);
    if (rc != PPK_OK) return rc;
    h->iter_host += 1;
  }
  int rc = hread_ctrl(h);
  if (rc != PPK_OK) return rc;
  h->iter_host = (long long)h->ctrl_host[ppk::CTRL_ITER];
  if (h->iter_host == start) h->cfl_valid = cfl_valid0;
  return PPK_OK;
}

int ppk_hydro_get_time(ppk_hydro_t *h, double *t, double *dt, int64_t *iteration) {
  if (!h) return hfail(h, PPK_ERR_INVALID_ARGUMENT, "NULL handle");
  int rc = hread_ctrl(h);
  if (rc != PPK_OK) return rc;
  if (t) *t = h->ctrl_host[ppk::CTRL_T];
  if (dt) *dt = h->ctrl_host[ppk::CTRL_DT];
  if (iteration) *iteration = (int64_t)h->ctrl_host[ppk::CTRL_ITER];
  return PPK_OK;
}

int ppk_hydro_set_time(ppk_hydro_t *h, double t, int64_t iteration) {
  if (!h) return hfail(h, PPK_ERR_INVALID_ARGUMENT, "NULL handle");
  if (iteration < 0) return hfail(h, PPK_ERR_INVALID_ARGUMENT, "iteration must be >= 0");
  int rc = hread_ctrl(h);
  if (rc != PPK_OK) return rc;
  if (((long long)iteration & 1) != (h->iter_host & 1)) std::swap(h->U[0], h->U[1]);
  h->cfl_valid = false;
  h->iter_host = iteration;
  h->ctrl_host[ppk::CTRL_T] = t;
  h->ctrl_host[ppk::CTRL_ITER] = (double)iteration;
  return hwrite_ctrl(h);
}

int ppk_hydro_finished(ppk_hydro_t *h, int *finished) {
  if (!h || !finished) return hfail(h, PPK_ERR_INVALID_ARGUMENT, "NULL argument");
  int rc = hread_ctrl(h);
  if (rc != PPK_OK) return rc;
  *finished = (h->ctrl_host[ppk::CTRL_T] >= (h->p.tEnd - 1e-14) ||
               h->ctrl_host[ppk::CTRL_ITER] >= (double)h->p.n_step_max)
                  ? 1
                  : 0;
  return PPK_OK;
}

int ppk_hydro_synchronize(ppk_hydro_t *h) {
  if (!h) return hfail(h, PPK_ERR_INVALID_ARGUMENT, "NULL handle");
  HHIP(h, hipStreamSynchronize(h->stream));
  return PPK_OK;
}

int ppk_hydro_timers_enable(ppk_hydro_t *h, uint32_t stage_mask) {
  if (!h) return hfail(h, PPK_ERR_INVALID_ARGUMENT, "NULL handle");
  HHIP(h, hipStreamSynchronize(h->stream));
  h->timer->set_mask(stage_mask);
  return PPK_OK;
}

int ppk_hydro_timers_read(ppk_hydro_t *h, int stage, double *total_ms, int64_t *launches) {
  if (!h) return hfail(h, PPK_ERR_INVALID_ARGUMENT, "NULL handle");
  if (stage < 0 || stage >= ppk::HS_COUNT) return hfail(h, PPK_ERR_INVALID_ARGUMENT, "bad stage");
  HHIP(h, h->timer->resolve());
  if (total_ms) *total_ms = h->timer->total(stage);
  if (launches) *launches = h->timer->count(stage);
  return PPK_OK;
}

int ppk_hydro_timer_count(void) { return ppk::HS_COUNT; }

const char *ppk_hydro_timer_name(int stage) {
  if (stage < 0 || stage >= ppk::HS_COUNT) return "";
  return ppk::kHydroStageNames[stage];
}

int ppk_hydro_save_vtk(ppk_hydro_t *h, const char *dir, const char *prefix, int istep,
                       int ascii) {
  if (!h || !dir || !prefix) return hfail(h, PPK_ERR_INVALID_ARGUMENT, "NULL argument");
  std::vector<double> host(h->values());
  int rc = ppk_hydro_download(h, host.data(), host.size());
  if (rc != PPK_OK) return rc;
  rc = ppk_mhd_write_vtk(&h->p, host.data(), dir, prefix, istep, ascii);
  if (rc != PPK_OK) return hfail(h, rc, ppk::thread_error());
  return PPK_OK;
}

int ppk_hydro_stream(ppk_hydro_t *h, void **stream) {
  if (!h || !stream) return hfail(h, PPK_ERR_INVALID_ARGUMENT, "NULL argument");
  *stream = (void *)h->stream;
  return PPK_OK;
}

}  // extern "C"
