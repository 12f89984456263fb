// VTK ImageData (.vti) output in the reference's format, src/utils/io/IO_VTK.cpp:23-207
// (2D) and :211-410 (3D): one CellData array per conservative variable over the interior cells,
// either "ascii" (the reference's outputVtkAscii = true) or appended raw binary (UInt64 byte
// count + the values, i fastest).  File name <dir>/<prefix>_<istep, 7 digits>.vti.  Header
// numbers go through a default std::ostream exactly as the reference writes them, so files are
// byte-identical to the reference's for the same state.
#include <cstdint>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <sstream>
#include <string>
#include <vector>

#include "capi_internal.hpp"
#include "ppkmhd_amd.h"

namespace {

// names of src/shared/SolverBase.cpp:48-55
const char *const kVarNames[8] = {"rho", "energy", "rho_vx", "rho_vy", "rho_vz", "bx", "by", "bz"};

bool big_endian() {
  const int one = 1;
  return *(const char *)&one == 0;
}

}  // namespace

extern "C" int ppk_mhd_write_vtk(const ppk_mhd_params *p, const double *U, const char *dir,
                                 const char *prefix, int istep, int ascii) {
  if (!p || !U || !dir || !prefix) {
    ppk::set_thread_error("ppk_mhd_write_vtk: NULL argument");
    return PPK_ERR_INVALID_ARGUMENT;
  }
  if (p->dim != 2 && p->dim != 3) {
    ppk::set_thread_error("ppk_mhd_write_vtk: dim must be 2 or 3");
    return PPK_ERR_INVALID_ARGUMENT;
  }
  const int gw = p->ghost_width, nx = p->nx, ny = p->ny, nz = p->dim == 3 ? p->nz : 1;
  // nbvar of the solver the parameters belong to: MHD (ghost width 3) 8, hydro (ghost width 2)
  // 4 in 2D and 5 in 3D (HydroParams::setup, src/shared/HydroParams.cpp:42-71)
  const int nbvar = gw == 2 ? (p->dim == 3 ? 5 : 4) : 8;
  const long long isize = nx + 2 * gw, jsize = ny + 2 * gw, ksize = p->dim == 3 ? nz + 2 * gw : 1;
  const long long vstride = isize * jsize * ksize;

  std::ostringstream step;
  step.width(7);
  step.fill('0');
  step << istep;
  const std::string filename = std::string(dir) + "/" + prefix + "_" + step.str() + ".vti";
  std::ofstream out(filename.c_str(), std::ios_base::out | std::ios_base::binary);
  if (!out) {
    ppk::set_thread_error("ppk_mhd_write_vtk: cannot open " + filename);
    return PPK_ERR_CONFIG;
  }

  if (ascii) out << "<?xml version=\"1.0\"?>\n";
  // the reference writes version 1.0 in 2D and 0.1 in 3D
  out << "<VTKFile type=\"ImageData\" version=\"" << (p->dim == 2 ? "1.0" : "0.1")
      << "\" byte_order=\"" << (big_endian() ? "BigEndian" : "LittleEndian")
      << "\" header_type=\"UInt64\">\n";
  if (p->dim == 2) {
    out << "  <ImageData WholeExtent=\"" << 0 << " " << nx << " " << 0 << " " << ny << " " << 0
        << " " << 0 << "\" " << "Origin=\"" << p->xmin << " " << p->ymin << " " << 0 << "\" "
        << "Spacing=\"" << p->dx << " " << p->dy << " " << 0.0 << "\">\n";
    out << "  <Piece Extent=\"" << 0 << " " << nx << " " << 0 << " " << ny << " " << 0 << " "
        << 0 << " " << "\">\n";
  } else {
    out << "  <ImageData WholeExtent=\"" << 0 << " " << nx << " " << 0 << " " << ny << " " << 0
        << " " << nz << "\" " << "Origin=\"" << p->xmin << " " << p->ymin << " " << p->zmin
        << "\" " << "Spacing=\"" << p->dx << " " << p->dy << " " << p->dz << "\">\n";
    out << "  <Piece Extent=\"" << 0 << " " << nx << " " << 0 << " " << ny << " " << 0 << " "
        << nz << " " << "\">\n";
  }
  out << "    <PointData>\n";
  out << "    </PointData>\n";

  auto interior = [&](int v, auto &&emit) {
    for (long long k = (p->dim == 3 ? gw : 0); k < (p->dim == 3 ? ksize - gw : 1); ++k)
      for (long long j = gw; j < jsize - gw; ++j)
        for (long long i = gw; i < isize - gw; ++i) emit(U[v * vstride + (k * jsize + j) * isize + i]);
  };

  if (ascii) {
    out << "    <CellData>\n";
    for (int v = 0; v < nbvar; ++v) {
      out << "    <DataArray type=\"Float64\" Name=\"" << kVarNames[v] << "\" format=\"ascii\" >\n";
      interior(v, [&](double x) { out << x << " "; });
      out << "\n    </DataArray>\n";
    }
    out << "    </CellData>\n";
    out << "  </Piece>\n";
    out << "  </ImageData>\n";
    out << "</VTKFile>\n";
  } else {
    const uint64_t words = (uint64_t)nx * ny * nz * sizeof(double);
    out << "    <CellData>" << std::endl;
    for (int v = 0; v < nbvar; ++v) {
      out << "     <DataArray type=\"Float64\" Name=\"" << kVarNames[v]
          << "\" format=\"appended\" offset=\"" << v * words + v * sizeof(uint64_t) << "\" />"
          << std::endl;
    }
    out << "    </CellData>" << std::endl;
    out << "  </Piece>" << std::endl;
    out << "  </ImageData>" << std::endl;
    out << "  <AppendedData encoding=\"raw\">" << std::endl;
    out << "_";
    std::vector<double> buf;
    buf.reserve((size_t)nx * ny * nz);
    for (int v = 0; v < nbvar; ++v) {
      out.write((const char *)&words, sizeof(uint64_t));
      buf.clear();
      interior(v, [&](double x) { buf.push_back(x); });
      out.write((const char *)buf.data(), (std::streamsize)(buf.size() * sizeof(double)));
    }
    out << "  </AppendedData>" << std::endl;
    out << "</VTKFile>" << std::endl;
  }
  if (!out) {
    ppk::set_thread_error("ppk_mhd_write_vtk: write failed on " + filename);
    return PPK_ERR_CONFIG;
  }
  return PPK_OK;
}
