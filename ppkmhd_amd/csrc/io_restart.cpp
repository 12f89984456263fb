// Restart files: the state with its ghost cells and the time-loop counters, so that a run can
// stop and continue with bit-identical results.
//
// The reference writes restart data with HDF5 (IO_HDF5.h:430-484: the fields, plus the root
// attributes "time step", "total time", nx, ny, nz and "ghost zone included") and reads them
// back in SolverMHDMuscl::init_restart (SolverMHDMuscl.h:613-643).  No HDF5 library is available
// here, so the same content goes into one raw little-endian file:
//
//   bytes   0..7    magic "PPKMHDRS"
//           8..39   int32 version (1), dim, nx, ny, nz, ghost_width, nbvar (8), ghost zone
//                   included (1)
//          40..55   int64 "time step" (the caller's output counter), int64 iteration
//          56..95   double "total time", gamma0, dx, dy, dz
//          96..127  zero
//         128..     nbvar x ksize x jsize x isize doubles, variable-major, i fastest (the
//                   device layout, ghosts included)
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "capi_internal.hpp"
#include "ppkmhd_amd.h"

namespace {

constexpr char kMagic[8] = {'P', 'P', 'K', 'M', 'H', 'D', 'R', 'S'};
constexpr int kHeader = 128;

struct Header {
  char magic[8];
  int32_t version, dim, nx, ny, nz, gw, nbvar, ghosts;
  int64_t time_step, iteration;
  double total_time, gamma0, dx, dy, dz;
  char pad[32];
};
static_assert(sizeof(Header) == kHeader, "restart header layout");

size_t cells_of(const ppk_mhd_params &p) {
  const size_t g = 2 * (size_t)p.ghost_width;
  return (p.nx + g) * (p.ny + g) * (p.dim == 3 ? p.nz + g : 1);
}

int err(int code, const std::string &m) {
  ppk::set_thread_error(m);
  return code;
}

}  // namespace

extern "C" {

int ppk_mhd_write_restart(const ppk_mhd_params *params, const double *U, double total_time,
                          int64_t time_step, int64_t iteration, const char *path) {
  if (!params || !U || !path) return err(PPK_ERR_INVALID_ARGUMENT, "NULL argument");
  Header h{};
  std::memcpy(h.magic, kMagic, 8);
  h.version = 1;
  h.dim = params->dim;
  h.nx = params->nx;
  h.ny = params->ny;
  h.nz = params->dim == 3 ? params->nz : 1;
  h.gw = params->ghost_width;
  h.nbvar = 8;
  h.ghosts = 1;
  h.time_step = time_step;
  h.iteration = iteration;
  h.total_time = total_time;
  h.gamma0 = params->gamma0;
  h.dx = params->dx;
  h.dy = params->dy;
  h.dz = params->dz;
  FILE *f = std::fopen(path, "wb");
  if (!f) return err(PPK_ERR_INVALID_ARGUMENT, std::string("cannot open ") + path);
  const size_t n = 8 * cells_of(*params);
  const bool ok = std::fwrite(&h, sizeof h, 1, f) == 1 && std::fwrite(U, sizeof(double), n, f) == n;
  const bool closed = std::fclose(f) == 0;
  if (!ok || !closed) return err(PPK_ERR_INVALID_ARGUMENT, std::string("write failed: ") + path);
  return PPK_OK;
}

int ppk_mhd_read_restart(const ppk_mhd_params *params, const char *path, double *U, size_t count,
                         double *total_time, int64_t *time_step, int64_t *iteration) {
  if (!params || !path || !U) return err(PPK_ERR_INVALID_ARGUMENT, "NULL argument");
  const size_t n = 8 * cells_of(*params);
  if (count != n) return err(PPK_ERR_INVALID_ARGUMENT, "count must be 8*isize*jsize*ksize");
  FILE *f = std::fopen(path, "rb");
  if (!f) return err(PPK_ERR_INVALID_ARGUMENT, std::string("cannot open ") + path);
  Header h{};
  if (std::fread(&h, sizeof h, 1, f) != 1 || std::memcmp(h.magic, kMagic, 8) != 0 ||
      h.version != 1) {
    std::fclose(f);
    return err(PPK_ERR_INVALID_ARGUMENT, std::string("not a restart file of this library: ") + path);
  }
  const int nz = params->dim == 3 ? params->nz : 1;
  if (h.dim != params->dim || h.nx != params->nx || h.ny != params->ny || h.nz != nz ||
      h.gw != params->ghost_width || h.nbvar != 8 || h.ghosts != 1) {
    std::fclose(f);
    char m[256];
    std::snprintf(m, sizeof m,
                  "restart file %s holds a %dD %dx%dx%d grid (ghost width %d), the solver "
                  "is %dD %dx%dx%d (ghost width %d)",
                  path, h.dim, h.nx, h.ny, h.nz, h.gw, params->dim, params->nx, params->ny, nz,
                  params->ghost_width);
    return err(PPK_ERR_INVALID_ARGUMENT, m);
  }
  const bool ok = std::fread(U, sizeof(double), n, f) == n;
  std::fclose(f);
  if (!ok) return err(PPK_ERR_INVALID_ARGUMENT, std::string("restart file truncated: ") + path);
  if (total_time) *total_time = h.total_time;
  if (time_step) *time_step = h.time_step;
  if (iteration) *iteration = h.iteration;
  return PPK_OK;
}

int ppk_mhd_save_state(ppk_mhd_t *h, const char *path, int64_t time_step) {
  if (!h || !path) return err(PPK_ERR_INVALID_ARGUMENT, "NULL argument");
  ppk_mhd_params p = ppk::handle_params(h);
  std::vector<double> U(8 * cells_of(p));
  double t = 0.0;
  int64_t it = 0;
  int rc = ppk_mhd_get_time(h, &t, nullptr, &it);
  if (rc == PPK_OK) rc = ppk_mhd_download(h, U.data(), U.size());
  if (rc != PPK_OK) return rc;
  rc = ppk_mhd_write_restart(&p, U.data(), t, time_step, it, path);
  if (rc != PPK_OK) ppk::set_handle_error(h, ppk::thread_error());
  return rc;
}

int ppk_mhd_load_state(ppk_mhd_t *h, const char *path, int64_t *time_step) {
  if (!h || !path) return err(PPK_ERR_INVALID_ARGUMENT, "NULL argument");
  ppk_mhd_params p = ppk::handle_params(h);
  std::vector<double> U(8 * cells_of(p));
  double t = 0.0;
  int64_t it = 0;
  int rc = ppk_mhd_read_restart(&p, path, U.data(), U.size(), &t, time_step, &it);
  if (rc != PPK_OK) {
    ppk::set_handle_error(h, ppk::thread_error());
    return rc;
  }
  // the counters first (their parity picks the buffer the state goes to), then the state
  rc = ppk_mhd_set_time(h, t, it);
  if (rc == PPK_OK) rc = ppk_mhd_upload(h, U.data(), U.size());
  return rc;
}

}  // extern "C"
