// Host run of the library's one-pass ghost fill (ghost_shell_element, kernels.hpp), built
// with AddressSanitizer / UndefinedBehaviorSanitizer.  The test compares it with the oracle's
// per-direction passes.
//
// Usage: ghost_fill <dir>
//   reads  dir/params.bin  ppk_mhd_params
//          dir/u0.bin      state (8 x cells doubles; ghosts arbitrary)
//   writes dir/u_shell.bin after the fill (dir/u_copy.bin: its copy target, which starts as
//                          NaN and must receive exactly the ghost values)
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "kernels.hpp"

using namespace ppk;

namespace {

bool read_file(const std::string &path, void *dst, size_t bytes) {
  FILE *f = std::fopen(path.c_str(), "rb");
  if (!f) return false;
  const size_t n = std::fread(dst, 1, bytes, f);
  std::fclose(f);
  return n == bytes;
}
bool write_file(const std::string &path, const std::vector<double> &v) {
  FILE *f = std::fopen(path.c_str(), "wb");
  if (!f) return false;
  const bool ok = std::fwrite(v.data(), sizeof(double), v.size(), f) == v.size();
  std::fclose(f);
  return ok;
}

}  // namespace

int main(int argc, char **argv) {
  if (argc != 2) {
    std::fprintf(stderr, "usage: ghost_fill <dir>\n");
    return 2;
  }
  const std::string dir = argv[1];
  ppk_mhd_params P;
  if (!read_file(dir + "/params.bin", &P, sizeof(P))) return 2;
  DevParams p{};
  p.dim = P.dim;
  p.nx = P.nx;
  p.ny = P.ny;
  p.nz = P.dim == 3 ? P.nz : 1;
  p.gw = P.ghost_width;
  p.isize = p.nx + 2 * p.gw;
  p.jsize = p.ny + 2 * p.gw;
  p.ksize = P.dim == 3 ? p.nz + 2 * p.gw : 1;
  for (int f = 0; f < 6; ++f) p.bc[f] = P.bc[f];
  p.plane = (long long)p.isize * p.jsize;
  p.vstride = p.plane * p.ksize;
  const size_t n = (size_t)p.vstride * 8;
  std::vector<double> U0(n);
  if (!read_file(dir + "/u0.bin", U0.data(), n * sizeof(double))) return 2;

  std::vector<double> B = U0, C(n, std::nan(""));
  for (long long t = 0; t < ghost_shell_cells(p); ++t)
    ghost_shell_element(B.data(), p, t, C.data());

  return write_file(dir + "/u_shell.bin", B) && write_file(dir + "/u_copy.bin", C) ? 0 : 2;
}
