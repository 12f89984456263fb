// Host emulation of the single-pass 3D MHD kernel (k_march3d): the kernel's own block program
// (march_block, march_body.hpp) with the threads of a block run one after the other between
// the block's synchronisation points, in the strict arithmetic.  Built for the CPU with
// AddressSanitizer and UndefinedBehaviorSanitizer, it checks the kernel's indexing and data
// flow without a GPU: every LDS slot and every per-thread register starts as NaN, so a value
// read before it is written shows up in the output, and every interior cell of the new state
// starts as NaN, so a cell no block writes shows up too.
//
// Usage: march_emu <dir>
//   reads  dir/params.bin  ppk_mhd_params (3D)
//          dir/u0.bin      state with its ghosts filled (8 x cells doubles)
//          dir/meta.txt    "dt blocks" (the number of blocks the grid's tile-planes are cut into)
//   writes dir/u1.bin      new state (ghosts: the input's, as the ghost-copy kernel writes them)
//          dir/cfl_cells.bin  the CFL term of each new interior cell (NaN elsewhere)
//          dir/cfl.txt     the maximum CFL term of the new interior cells (%.17g)
#include <cmath>
#include <cstdio>
#include <cstring>
#include <type_traits>
#include <vector>

#include "march_body.hpp"

using namespace ppk;
using namespace ppk::PPK_NS;
using namespace ppk::PPK_NS::marchk;

namespace {

template <class T>
struct HostExec {
  std::vector<T> &ts;
  const DevParams &p;
  std::vector<double> &cells;  // CFL term per cell (NaN where no block computed one)
  int tid(const T &t) const { return (int)(&t - ts.data()); }
  template <class F>
  void each(F &&f) {
    for (auto &t : ts) f(t);
  }
  void cell_cfl(const T &t, int k, double term) { cells[cidx(p, t.i, t.j, k)] = term; }
};

DevParams dev_params(const ppk_mhd_params &p) {
  DevParams d{};
  d.dim = p.dim;
  d.nx = p.nx;
  d.ny = p.ny;
  d.nz = p.nz;
  d.gw = p.ghost_width;
  d.isize = p.nx + 2 * d.gw;
  d.jsize = p.ny + 2 * d.gw;
  d.ksize = p.nz + 2 * d.gw;
  d.riemann = p.riemann;
  for (int f = 0; f < 6; ++f) d.bc[f] = p.bc[f];
  d.slope12 = (p.slope_type == 1.0 || p.slope_type == 2.0) ? 1 : 0;
  d.plane = (long long)d.isize * d.jsize;
  d.vstride = d.plane * d.ksize;
  d.dx = p.dx;
  d.dy = p.dy;
  d.dz = p.dz;
  d.gamma0 = p.gamma0;
  d.slope_type = p.slope_type;
  d.smallr = p.smallr;
  d.smallc = p.smallc;
  d.smallp = p.smallp;
  d.cIso = p.cIso;
  d.cfl = p.cfl;
  return d;
}

template <int RS, int ST>
double run(const DevParams &p, int nblk, const double *Uin, double *Uout, double dt,
           std::vector<double> &cells) {
  using T = MarchThread<RS, ST>;
  const MarchGeom g = march_geom(p, nblk);
  std::vector<double> lds(kMarchLds);
  std::vector<T> ts(kMN);
  double cfl = 0.0;
  for (long long L = 0; L < g.blocks(); ++L) {
    for (double &x : lds) x = std::nan("");
    std::memset(static_cast<void *>(ts.data()), 0xff, sizeof(T) * ts.size());
    HostExec<T> X{ts, p, cells};
    march_block<RS, ST>(X, Uin, Uout, p, g, (int)L, dt, true, lds.data());
    for (const T &t : ts) {
      if (std::isnan(t.cfl)) return std::nan("");
      cfl = std::fmax(cfl, t.cfl);
    }
  }
  return cfl;
}

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
    std::fprintf(stderr, "usage: march_emu <dir>\n");
    return 2;
  }
  const std::string dir = argv[1];
  ppk_mhd_params P;
  if (!read_file(dir + "/params.bin", &P, sizeof(P)) || P.dim != 3) {
    std::fprintf(stderr, "bad params.bin\n");
    return 2;
  }
  const DevParams p = dev_params(P);
  const size_t n = (size_t)p.vstride * 8;
  std::vector<double> U0(n), U1(n);
  if (!read_file(dir + "/u0.bin", U0.data(), n * sizeof(double))) {
    std::fprintf(stderr, "bad u0.bin\n");
    return 2;
  }
  double dt = 0.0;
  int nblk = 0;
  {
    FILE *f = std::fopen((dir + "/meta.txt").c_str(), "r");
    if (!f || std::fscanf(f, "%lf %d", &dt, &nblk) != 2) {
      std::fprintf(stderr, "bad meta.txt\n");
      return 2;
    }
    std::fclose(f);
  }
  // ghosts: the input's; interior: NaN until a block writes it
  U1 = U0;
  for (int v = 0; v < 8; ++v)
    for (int k = p.gw; k < p.ksize - p.gw; ++k)
      for (int j = p.gw; j < p.jsize - p.gw; ++j)
        for (int i = p.gw; i < p.isize - p.gw; ++i)
          U1[(size_t)v * p.vstride + cidx(p, i, j, k)] = std::nan("");
  std::vector<double> cells((size_t)p.vstride, std::nan(""));
  // the kernel's instantiations: minmod slopes with the factor fixed, others read it
  auto go = [&](auto st) {
    constexpr int ST = decltype(st)::value;
    if (p.riemann == PPK_RIEMANN_HLLD)
      return run<PPK_RIEMANN_HLLD, ST>(p, nblk, U0.data(), U1.data(), dt, cells);
    if (p.riemann == PPK_RIEMANN_HLL)
      return run<PPK_RIEMANN_HLL, ST>(p, nblk, U0.data(), U1.data(), dt, cells);
    return run<PPK_RIEMANN_LLF, ST>(p, nblk, U0.data(), U1.data(), dt, cells);
  };
  const double cfl = p.slope_type == 1.0 ? go(std::integral_constant<int, 1>{})
                                         : go(std::integral_constant<int, 0>{});
  FILE *f = std::fopen((dir + "/u1.bin").c_str(), "wb");
  if (!f || std::fwrite(U1.data(), sizeof(double), n, f) != n) return 2;
  std::fclose(f);
  f = std::fopen((dir + "/cfl_cells.bin").c_str(), "wb");
  if (!f || std::fwrite(cells.data(), sizeof(double), cells.size(), f) != cells.size()) return 2;
  std::fclose(f);
  f = std::fopen((dir + "/cfl.txt").c_str(), "w");
  if (!f) return 2;
  std::fprintf(f, "%.17g\n", cfl);
  std::fclose(f);
  return 0;
}
