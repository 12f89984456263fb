// A compiled C++ caller of the library's C ABI, written the way the reference-side binding
// (a SolverBase subclass registered with SolverFactory::registerSolver) calls it:
//
//  - the solver parameters are copied field by field from a HydroParams-shaped struct, then
//    ppk_mhd_params_init derives dx, smallp, smallpp and gamma6 (HydroParams::init,
//    src/shared/HydroParams.cpp:421-453);
//  - the problem parameters come through ppk_mhd_params_from_lookup from a ConfigMap-shaped
//    key/value store (the reference builds BlastParams(configMap), ... from the same map);
//  - the time loop is src/main.cpp:143-166 (while (!finished()) next_iteration()) with the
//    TIMER_TOTAL / TIMER_NUM_SCHEME / TIMER_DT / TIMER_BOUNDARIES / TIMER_IO accumulators that
//    print_solver_monitoring_info reads (src/shared/solver_utils.h:16-60), and its "Perf" line.
//
// Usage: abi_caller errors   parameter and error paths only (no GPU needed)
//        abi_caller run      the above, then Orszag-Tang 2D and blast 3D runs on the GPU
// Exit status 0 when every check holds; each failed check prints a line naming it.
#include <chrono>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "ppkmhd_amd.h"

namespace {

int g_failures = 0;
#define CHECK(cond)                                                        \
  do {                                                                     \
    if (!(cond)) {                                                         \
      std::printf("FAILED %s:%d: %s\n", __FILE__, __LINE__, #cond);        \
      ++g_failures;                                                        \
    }                                                                      \
  } while (0)

// ---- the reference's types, as far as the binding reads them -----------------------------
// HydroParams / HydroSettings fields (src/shared/HydroParams.h:28-234); enum values of
// RiemannSolverType and BoundaryConditionType (src/shared/enums.h:136-155) equal the
// library's ppk_riemann / ppk_bc values.
struct HydroSettings {
  double gamma0 = 1.4, cfl = 0.5, slope_type = 1.0, smallr = 1e-10, smallc = 1e-10, cIso = 0.0;
};
struct HydroParams {
  int nStepmax = 1000;
  double tEnd = 0.0;
  int nx = 1, ny = 1, nz = 1, ghostWidth = 3;
  int dimType = 2;  // TWO_D / THREE_D
  double xmin = 0, xmax = 1, ymin = 0, ymax = 1, zmin = 0, zmax = 1;
  int boundary_type_xmin = 1, boundary_type_xmax = 1, boundary_type_ymin = 1;
  int boundary_type_ymax = 1, boundary_type_zmin = 1, boundary_type_zmax = 1;
  int riemannSolverType = 4;
  int implementationVersion = 0;
  HydroSettings settings;
};
// ConfigMap: INIReader semantics, case-insensitive "section.name" keys
struct ConfigMap {
  std::map<std::string, std::string> kv;
  static std::string key(const char *s, const char *n) {
    std::string k = std::string(s) + "." + n;
    for (auto &c : k) c = (char)std::tolower((unsigned char)c);
    return k;
  }
  void set(const char *s, const char *n, const std::string &v) { kv[key(s, n)] = v; }
};

// the ppk_ini_lookup a binding passes: the raw string of [section] name, or 0 if absent
int config_lookup(void *ctx, const char *section, const char *name, char *value, int size) {
  const ConfigMap &cm = *static_cast<const ConfigMap *>(ctx);
  auto it = cm.kv.find(ConfigMap::key(section, name));
  if (it == cm.kv.end()) return 0;
  std::snprintf(value, (size_t)size, "%s", it->second.c_str());
  return 1;
}

// ppk_mhd_params from the HydroParams fields, one to one
ppk_mhd_params from_hydro_params(const HydroParams &hp, int precision) {
  ppk_mhd_params p;
  std::memset(&p, 0, sizeof p);
  p.dim = hp.dimType;
  p.nx = hp.nx;
  p.ny = hp.ny;
  p.nz = hp.dimType == 3 ? hp.nz : 1;
  p.ghost_width = hp.ghostWidth;
  p.riemann = hp.riemannSolverType;
  const int bc[6] = {hp.boundary_type_xmin, hp.boundary_type_xmax, hp.boundary_type_ymin,
                     hp.boundary_type_ymax, hp.boundary_type_zmin, hp.boundary_type_zmax};
  for (int f = 0; f < 6; ++f) p.bc[f] = bc[f];
  p.n_step_max = hp.nStepmax;
  p.precision = precision;
  p.kernel_path = PPK_PATH_AUTO;
  p.implementation_version = hp.implementationVersion;
  p.xmin = hp.xmin; p.xmax = hp.xmax;
  p.ymin = hp.ymin; p.ymax = hp.ymax;
  p.zmin = hp.zmin; p.zmax = hp.zmax;
  p.gamma0 = hp.settings.gamma0;
  p.cfl = hp.settings.cfl;
  p.slope_type = hp.settings.slope_type;
  p.smallr = hp.settings.smallr;
  p.smallc = hp.settings.smallc;
  p.cIso = hp.settings.cIso;
  p.tEnd = hp.tEnd;
  ppk_mhd_params_init(&p);
  return p;
}

// SolverBase's timers (OpenMPTimer: start / stop accumulate wall time)
enum { TIMER_TOTAL = 0, TIMER_IO, TIMER_DT, TIMER_BOUNDARIES, TIMER_NUM_SCHEME, TIMER_COUNT };
struct Timer {
  double total = 0.0;
  std::chrono::steady_clock::time_point t0;
  void start() { t0 = std::chrono::steady_clock::now(); }
  void stop() {
    total += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
  double elapsed() const { return total; }
};

// the binding: SolverBase-shaped state, C-ABI calls underneath
struct SolverMHDMusclHip {
  ppk_mhd_t *h = nullptr;
  double m_t = 0.0, m_dt = 0.0;
  int m_iteration = 0;
  long long m_nCells = 0;
  Timer timers[TIMER_COUNT];
  int status = PPK_OK;

  SolverMHDMusclHip(const HydroParams &hp, ConfigMap &cm, int precision) {
    ppk_mhd_params p = from_hydro_params(hp, precision);
    ppk_mhd_params q;
    ppk_mhd_problem prob;
    char name[64];
    // problem parameters from the ConfigMap (the solver parameters stay the HydroParams')
    status = ppk_mhd_params_from_lookup(config_lookup, &cm, &q, &prob, name, (int)sizeof name);
    if (status == PPK_OK) status = ppk_mhd_create(&p, &h);
    if (status == PPK_OK) status = ppk_mhd_init_problem(h, &prob);
    m_nCells = (long long)(p.nx + 2 * p.ghost_width) * (p.ny + 2 * p.ghost_width) *
               (p.dim == 3 ? p.nz + 2 * p.ghost_width : 1);
    if (status == PPK_OK) sync_time();
  }
  ~SolverMHDMusclHip() { ppk_mhd_destroy(h); }

  void sync_time() {
    int64_t it = 0;
    status = ppk_mhd_get_time(h, &m_t, &m_dt, &it);
    m_iteration = (int)it;
  }
  int finished() {
    int f = 1;
    if (ppk_mhd_finished(h, &f) != PPK_OK) return 1;
    return f;
  }
  // SolverBase::next_iteration: the device does boundaries, dt, the scheme, ++iteration and
  // t += dt; the timers bracket the calls the way godunov_unsplit_impl brackets its stages
  // (boundaries and dt are launched inside the step, so they count under NUM_SCHEME)
  void next_iteration() {
    timers[TIMER_NUM_SCHEME].start();
    status = ppk_mhd_step(h, 0.0, PPK_DT_CFL);
    if (status == PPK_OK) status = ppk_mhd_synchronize(h);
    timers[TIMER_NUM_SCHEME].stop();
    if (status == PPK_OK) sync_time();
  }
  double compute_dt_local() {
    timers[TIMER_DT].start();
    double dt = 0.0;
    status = ppk_mhd_compute_dt(h, &dt);
    timers[TIMER_DT].stop();
    return dt;
  }
};

// print_solver_monitoring_info (src/shared/solver_utils.h:16-60)
double print_monitoring(SolverMHDMusclHip &s) {
  const double t_tot = s.timers[TIMER_TOTAL].elapsed();
  const double t_comp = s.timers[TIMER_NUM_SCHEME].elapsed();
  std::printf("total       time : %5.3f secondes\n", t_tot);
  std::printf("godunov     time : %5.3f secondes %5.2f%%\n", t_comp, 100 * t_comp / t_tot);
  const double perf = s.m_iteration * (double)s.m_nCells / t_tot * 1e-6;
  std::printf("Perf             : %5.3f number of Mcell-updates/s\n", perf);
  return perf;
}

ConfigMap orszag_tang_config() {
  ConfigMap cm;
  cm.set("run", "solver_name", "MHD_Muscl_2D");
  cm.set("hydro", "problem", "orszag_tang");
  return cm;
}

void check_error_paths() {
  HydroParams hp;
  hp.nx = hp.ny = 32;
  hp.settings.gamma0 = 5.0 / 3.0;
  ppk_mhd_t *h = nullptr;
  // invalid dimension: refused with a message, no handle
  ppk_mhd_params p = from_hydro_params(hp, PPK_FP64);
  p.dim = 4;
  CHECK(ppk_mhd_create(&p, &h) == PPK_ERR_INVALID_ARGUMENT);
  CHECK(h == nullptr);
  CHECK(std::strlen(ppk_mhd_last_error(nullptr)) > 0);
  // implementation version 2 (SolverMHDMuscl.cpp:430-456) is not provided
  hp.implementationVersion = 2;
  p = from_hydro_params(hp, PPK_FP64);
  CHECK(ppk_mhd_create(&p, &h) == PPK_ERR_UNSUPPORTED);
  CHECK(std::strstr(ppk_mhd_last_error(nullptr), "implementationVersion") != nullptr);
  hp.implementationVersion = 0;
  // a solver the library does not provide (SolverFactory aborts on it)
  ConfigMap bad;
  bad.set("run", "solver_name", "Hydro_SDM_2D_degree2");
  ppk_mhd_params q;
  ppk_mhd_problem prob;
  char name[64];
  CHECK(ppk_mhd_params_from_lookup(config_lookup, &bad, &q, &prob, name, 64) == PPK_ERR_CONFIG);
  CHECK(std::string(name) == "Hydro_SDM_2D_degree2");
  // the lookup gives what the .ini text gives
  ConfigMap cm = orszag_tang_config();
  cm.set("mesh", "nx", "64");
  cm.set("Hydro", "GAMMA0", "1.666666666667");
  CHECK(ppk_mhd_params_from_lookup(config_lookup, &cm, &q, &prob, name, 64) == PPK_OK);
  CHECK(q.nx == 64 && q.dim == 2 && prob.kind == PPK_PROBLEM_ORSZAG_TANG);
  CHECK(q.gamma0 == (double)1.666666666667f);  // reals through strtof, as ConfigMap::getFloat
  // calls on a NULL handle fail cleanly
  CHECK(ppk_mhd_step(nullptr, 0.0, PPK_DT_CFL) == PPK_ERR_INVALID_ARGUMENT);
  double dt = 0.0;
  CHECK(ppk_mhd_compute_dt(nullptr, &dt) == PPK_ERR_INVALID_ARGUMENT);
  ppk_mhd_destroy(nullptr);
}

void check_runs() {
  // Orszag-Tang 2D 64^2 to t = 0.05 through the reference's main loop
  HydroParams hp;
  hp.nx = hp.ny = 64;
  hp.tEnd = 0.05;
  hp.nStepmax = 100000;
  hp.settings.gamma0 = 5.0 / 3.0;
  hp.settings.cfl = 0.8;
  hp.boundary_type_xmin = hp.boundary_type_xmax = 3;
  hp.boundary_type_ymin = hp.boundary_type_ymax = 3;
  ConfigMap cm = orszag_tang_config();
  {
    SolverMHDMusclHip s(hp, cm, PPK_FP64);
    CHECK(s.status == PPK_OK);
    if (s.status != PPK_OK) {
      std::printf("create: %s\n", ppk_mhd_last_error(s.h));
      return;
    }
    const double dt0 = s.compute_dt_local();
    CHECK(s.status == PPK_OK && dt0 > 0.0 && dt0 < 0.05);
    s.timers[TIMER_TOTAL].start();
    while (!s.finished()) {
      s.next_iteration();
      if (s.status != PPK_OK) break;
    }
    s.timers[TIMER_TOTAL].stop();
    CHECK(s.status == PPK_OK);
    CHECK(s.m_iteration > 10 && std::fabs(s.m_t - hp.tEnd) < 1e-15);
    std::vector<double> U((size_t)8 * s.m_nCells);
    CHECK(ppk_mhd_download(s.h, U.data(), U.size()) == PPK_OK);
    bool finite = true;
    for (double v : U) finite = finite && std::isfinite(v);
    CHECK(finite);
    // density of the first interior cell stays near 25/(36 pi)
    const int isize = hp.nx + 6;
    CHECK(std::fabs(U[3 * isize + 3] - 25.0 / (36.0 * M_PI)) < 0.05);
    CHECK(print_monitoring(s) > 0.0);
    // a wrong-size download is refused and leaves a message on the handle
    CHECK(ppk_mhd_download(s.h, U.data(), U.size() - 1) == PPK_ERR_INVALID_ARGUMENT);
    CHECK(std::strlen(ppk_mhd_last_error(s.h)) > 0);
  }
  // blast 3D 16^3 (config defaults from the ConfigMap's [blast] section), 3 steps
  HydroParams hb;
  hb.dimType = 3;
  hb.nx = hb.ny = hb.nz = 16;
  hb.tEnd = 1.0;
  hb.nStepmax = 3;
  hb.settings.gamma0 = 5.0 / 3.0;
  hb.settings.cfl = 0.8;
  hb.boundary_type_xmin = hb.boundary_type_xmax = hb.boundary_type_ymin = 3;
  hb.boundary_type_ymax = hb.boundary_type_zmin = hb.boundary_type_zmax = 3;
  ConfigMap cb;
  cb.set("run", "solver_name", "MHD_Muscl_3D");
  cb.set("hydro", "problem", "blast");
  cb.set("blast", "radius", "0.2");
  cb.set("blast", "density_out", "1.0");
  SolverMHDMusclHip s(hb, cb, PPK_FP64_STRICT);
  CHECK(s.status == PPK_OK);
  if (s.status != PPK_OK) return;
  CHECK(ppk_mhd_run(s.h, 10, 0.0, PPK_DT_CFL) == PPK_OK);  // stops at nStepmax
  s.sync_time();
  CHECK(s.m_iteration == 3 && s.finished());
  double divb = 1.0;
  CHECK(ppk_mhd_divb_max(s.h, &divb) == PPK_OK && divb < 1e-12 * 16);
}

}  // namespace

int main(int argc, char **argv) {
  const std::string mode = argc > 1 ? argv[1] : "errors";
  CHECK(ppk_mhd_params_size() == (int)sizeof(ppk_mhd_params));
  CHECK(ppk_mhd_problem_size() == (int)sizeof(ppk_mhd_problem));
  check_error_paths();
  if (mode == "run") {
    CHECK(ppk_mhd_device_count() > 0);
    check_runs();
  }
  std::printf("%s: %d failed checks\n", mode.c_str(), g_failures);
  return g_failures == 0 ? 0 : 1;
}
