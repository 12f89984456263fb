// Driver of the CPU oracles for their AddressSanitizer / UndefinedBehaviorSanitizer builds:
// oracle/mhd_oracle.cpp (default) or, built with -DORACLE_HYDRO, oracle/hydro_oracle.cpp,
// compiled together with this file for the host.  It runs a problem for a number of CFL steps
// through the oracle's own C entry points, so that every stage of the restated step (ghost
// fill, primitives, dt, trace, solvers, EMFs, update, CT, initial conditions) executes under
// the sanitizers; the test compares the result with the regular oracle library bit for bit.
//
// Usage: oracle_asan <dir>
//   reads  dir/params.bin   ppk_mhd_params
//          dir/problem.bin  ppk_mhd_problem
//          dir/meta.txt     "steps"
//   writes dir/u.bin        state after the steps (8 or 5 x cells doubles, ghosts included)
//          dir/time.txt     "t dt iteration" (%.17g %.17g %lld)
#include <cstdio>
#include <string>
#include <vector>

#include "ppkmhd_amd.h"

extern "C" {
#ifdef ORACLE_HYDRO
typedef struct hora_solver hora_solver;
hora_solver *hora_create(const ppk_mhd_params *p);
void hora_destroy(hora_solver *s);
void hora_init_problem(hora_solver *s, const ppk_mhd_problem *pr);
void hora_get_state(hora_solver *s, double *U);
void hora_step(hora_solver *s, double dt, int mode);
void hora_get_time(hora_solver *s, double *t, double *dt, long long *it);
#else
typedef struct ora_solver ora_solver;
ora_solver *ora_create(const ppk_mhd_params *p);
void ora_destroy(ora_solver *s);
void ora_init_problem(ora_solver *s, const ppk_mhd_problem *prob);
void ora_get_state(const ora_solver *s, double *U);
double ora_step(ora_solver *s, double dt, int dt_mode);
void ora_get_time(const ora_solver *s, double *t, double *dt, long long *iteration);
#endif
}

namespace {

bool read_file(const std::string &path, void *dst, size_t bytes) {
  FILE *f = std::fopen(path.c_str(), "rb");
  if (!f) return false;
  const size_t n = std::fread(dst, 1, bytes, f);
  std::fclose(f);
  return n == bytes;
}

}  // namespace

int main(int argc, char **argv) {
  if (argc != 2) {
    std::fprintf(stderr, "usage: oracle_asan <dir>\n");
    return 2;
  }
  const std::string dir = argv[1];
  ppk_mhd_params p;
  ppk_mhd_problem pr;
  if (!read_file(dir + "/params.bin", &p, sizeof(p)) ||
      !read_file(dir + "/problem.bin", &pr, sizeof(pr))) {
    std::fprintf(stderr, "bad params.bin / problem.bin\n");
    return 2;
  }
  int steps = 0;
  {
    FILE *f = std::fopen((dir + "/meta.txt").c_str(), "r");
    if (!f || std::fscanf(f, "%d", &steps) != 1) {
      std::fprintf(stderr, "bad meta.txt\n");
      return 2;
    }
    std::fclose(f);
  }
  const int gw = p.ghost_width;
  const size_t cells = (size_t)(p.nx + 2 * gw) * (p.ny + 2 * gw) * (p.dim == 3 ? p.nz + 2 * gw : 1);
  double t = 0.0, dt = 0.0;
  long long it = 0;
#ifdef ORACLE_HYDRO
  std::vector<double> U(cells * (p.dim == 3 ? 5 : 4));
  hora_solver *s = hora_create(&p);
  hora_init_problem(s, &pr);
  for (int n = 0; n < steps; ++n) hora_step(s, 0.0, PPK_DT_CFL);
  hora_get_state(s, U.data());
  hora_get_time(s, &t, &dt, &it);
  hora_destroy(s);
#else
  std::vector<double> U(cells * 8);
  ora_solver *s = ora_create(&p);
  ora_init_problem(s, &pr);
  for (int n = 0; n < steps; ++n) ora_step(s, 0.0, PPK_DT_CFL);
  ora_get_state(s, U.data());
  ora_get_time(s, &t, &dt, &it);
  ora_destroy(s);
#endif
  FILE *f = std::fopen((dir + "/u.bin").c_str(), "wb");
  if (!f || std::fwrite(U.data(), sizeof(double), U.size(), f) != U.size()) return 2;
  std::fclose(f);
  f = std::fopen((dir + "/time.txt").c_str(), "w");
  if (!f) return 2;
  std::fprintf(f, "%.17g %.17g %lld\n", t, dt, it);
  std::fclose(f);
  return 0;
}
