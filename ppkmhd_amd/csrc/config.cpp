// .ini parameters -> ppk_mhd_params / ppk_mhd_problem, with the reference's semantics:
//  - INI syntax of the vendored inih reader: [section], name = value (or name: value),
//    full-line comments starting with ';' or '#', inline comments ';' after whitespace,
//    indented continuation lines replace the previous value
//    (reference src/utils/config/inih/ini.cpp);
//  - section and key lookups are case-insensitive (INIReader::makeKey lowercases,
//    src/utils/config/inih/INIReader.cpp:100-105);
//  - reals go through strtof and come back as float (ConfigMap::getFloat,
//    src/utils/config/ConfigMap.cpp:37-46); integers through strtol(.., 0);
//    booleans accept 1/yes/true/on and 0/no/false/off (ConfigMap::getBool :62-80);
//  - defaults and derived values of HydroParams::setup / init
//    (src/shared/HydroParams.cpp:28-217, :421-453) and of the problem parameter structs
//    (src/shared/problems/*Params.h).
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>

#include "capi_internal.hpp"
#include "ppkmhd_amd.h"

namespace ppk {
namespace {

std::string trim(const std::string &s) {
  size_t a = 0, b = s.size();
  while (a < b && std::isspace((unsigned char)s[a])) ++a;
  while (b > a && std::isspace((unsigned char)s[b - 1])) --b;
  return s.substr(a, b - a);
}

std::string lower(std::string s) {
  for (auto &c : s) c = (char)std::tolower((unsigned char)c);
  return s;
}

// key lookups with the reference's typed accessors on top of one raw string lookup
class IniMap {
 public:
  virtual ~IniMap() {}
  // the raw value of [s] n, if the key exists
  virtual bool raw(const std::string &s, const std::string &n, std::string &out) const = 0;

  std::string get_string(const std::string &s, const std::string &n, const std::string &d) const {
    std::string v;
    return raw(s, n, v) ? v : d;
  }
  long get_integer(const std::string &s, const std::string &n, long d) const {
    std::string v = get_string(s, n, "");
    const char *c = v.c_str();
    char *end = nullptr;
    long r = std::strtol(c, &end, 0);
    return end > c ? r : d;
  }
  float get_float(const std::string &s, const std::string &n, float d) const {
    std::string v = get_string(s, n, "");
    const char *c = v.c_str();
    char *end = nullptr;
    float r = std::strtof(c, &end);
    return end > c ? r : d;
  }
  bool get_bool(const std::string &s, const std::string &n, bool d) const {
    std::string v = get_string(s, n, "");
    bool r = d;
    if (v == "1" || v == "yes" || v == "true" || v == "on") r = true;
    if (v == "0" || v == "no" || v == "false" || v == "off") r = false;
    if (v.empty()) r = d;
    return r;
  }
};

// the text of an .ini file, parsed as the vendored inih reader does
class IniText : public IniMap {
 public:
  explicit IniText(const std::string &text) {
    std::istringstream in(text);
    std::string line, section, prev_name;
    while (std::getline(in, line)) {
      if (!line.empty() && line.back() == '\r') line.pop_back();
      bool indented = !line.empty() && std::isspace((unsigned char)line[0]);
      std::string t = trim(line);
      if (t.empty()) continue;
      if (t[0] == ';' || t[0] == '#') continue;
      if (indented && !prev_name.empty() && t[0] != '[') {
        values_[key(section, prev_name)] = strip_inline(t);  // continuation line
        continue;
      }
      if (t[0] == '[') {
        size_t e = t.find(']');
        if (e != std::string::npos) section = trim(t.substr(1, e - 1));
        prev_name.clear();
        continue;
      }
      size_t e = t.find_first_of("=:");
      if (e == std::string::npos) continue;
      std::string name = trim(t.substr(0, e));
      std::string value = strip_inline(trim(t.substr(e + 1)));
      values_[key(section, name)] = value;
      prev_name = name;
    }
  }
  bool raw(const std::string &s, const std::string &n, std::string &out) const override {
    auto it = values_.find(key(s, n));
    if (it == values_.end()) return false;
    out = it->second;
    return true;
  }

 private:
  static std::string key(const std::string &s, const std::string &n) { return lower(s + "." + n); }
  static std::string strip_inline(const std::string &v) {
    // ';' starts a comment only after whitespace (inih find_char_or_comment)
    bool ws = false;
    for (size_t i = 0; i < v.size(); ++i) {
      if (ws && v[i] == ';') return trim(v.substr(0, i));
      ws = std::isspace((unsigned char)v[i]) != 0;
    }
    return v;
  }
  std::map<std::string, std::string> values_;
};

// a caller's key lookup (the reference's ConfigMap behind a callback)
class IniCallback : public IniMap {
 public:
  IniCallback(ppk_ini_lookup fn, void *ctx) : fn_(fn), ctx_(ctx) {}
  bool raw(const std::string &s, const std::string &n, std::string &out) const override {
    char buf[1024];
    buf[0] = 0;
    if (!fn_(ctx_, s.c_str(), n.c_str(), buf, (int)sizeof buf)) return false;
    buf[sizeof buf - 1] = 0;
    out = buf;
    return true;
  }

 private:
  ppk_ini_lookup fn_;
  void *ctx_;
};

int fill_from_ini(const IniMap &m, ppk_mhd_params *p, ppk_mhd_problem *pr, char *solver_name,
                  int name_len) {
  ppk_mhd_params P{};
  // [run]
  P.n_step_max = (int)m.get_integer("run", "nstepmax", 1000);
  P.tEnd = m.get_float("run", "tend", 0.0f);
  std::string sname = m.get_string("run", "solver_name", "unknown");
  if (solver_name && name_len > 0) {
    std::snprintf(solver_name, (size_t)name_len, "%s", sname.c_str());
  }
  // HydroParams::setup, src/shared/HydroParams.cpp:42-71: solver name -> dim, ghostWidth
  P.ghost_width = 3;
  if (sname == "MHD_Muscl_2D") {
    P.dim = 2;
  } else if (sname == "MHD_Muscl_3D") {
    P.dim = 3;
  } else if (sname == "Hydro_Muscl_2D") {
    P.dim = 2;
    P.ghost_width = 2;
  } else if (sname == "Hydro_Muscl_3D") {
    P.dim = 3;
    P.ghost_width = 2;
  } else {
    P.dim = 0;
  }
  // [mesh]
  P.nx = (int)m.get_integer("mesh", "nx", 1);
  P.ny = (int)m.get_integer("mesh", "ny", 1);
  P.nz = (int)m.get_integer("mesh", "nz", 1);
  P.xmin = m.get_float("mesh", "xmin", 0.0f);
  P.ymin = m.get_float("mesh", "ymin", 0.0f);
  P.zmin = m.get_float("mesh", "zmin", 0.0f);
  P.xmax = m.get_float("mesh", "xmax", 1.0f);
  P.ymax = m.get_float("mesh", "ymax", 1.0f);
  P.zmax = m.get_float("mesh", "zmax", 1.0f);
  const char *bcn[6] = {"boundary_type_xmin", "boundary_type_xmax", "boundary_type_ymin",
                        "boundary_type_ymax", "boundary_type_zmin", "boundary_type_zmax"};
  for (int f = 0; f < 6; ++f) P.bc[f] = (int)m.get_integer("mesh", bcn[f], PPK_BC_DIRICHLET);
  // [hydro]
  P.gamma0 = m.get_float("hydro", "gamma0", 1.4f);
  P.cfl = m.get_float("hydro", "cfl", 0.5f);
  P.slope_type = m.get_float("hydro", "slope_type", 1.0f);
  P.smallc = m.get_float("hydro", "smallc", 1e-10f);
  P.smallr = m.get_float("hydro", "smallr", 1e-10f);
  P.cIso = 0.0;
  std::string rs = m.get_string("hydro", "riemann", "approx");
  if (rs == "approx") P.riemann = PPK_RIEMANN_APPROX;
  else if (rs == "llf") P.riemann = PPK_RIEMANN_LLF;
  else if (rs == "hll") P.riemann = PPK_RIEMANN_HLL;
  else if (rs == "hllc") P.riemann = PPK_RIEMANN_HLLC;
  else if (rs == "hlld") P.riemann = PPK_RIEMANN_HLLD;
  else P.riemann = PPK_RIEMANN_APPROX;
  // [other] extensions of this library (not in the reference)
  std::string prec = lower(m.get_string("other", "precision", "fp64"));
  P.precision = prec == "fp64_strict" ? PPK_FP64_STRICT
                : (prec == "fp32_bulk" ? PPK_FP32_BULK : PPK_FP64);
  std::string path = lower(m.get_string("other", "kernel_path", "auto"));
  P.kernel_path = path == "staged"           ? PPK_PATH_STAGED
                  : path == "fused"          ? PPK_PATH_FUSED
                  : path == "fused_records" ? PPK_PATH_FUSED_RECORDS
                                             : PPK_PATH_AUTO;
  P.implementation_version = (int)m.get_integer("OTHER", "implementationVersion", 0);
  // SolverBase: gravity enabled when [gravity] static or point (SolverBase.cpp:138-141); the
  // point-mass field is the disk problem's, which is not provided, so either key gives the
  // uniform field of the problem (zero for problems that set none)
  P.gravity_static =
      (m.get_bool("gravity", "static", false) || m.get_bool("gravity", "point", false)) ? 1 : 0;
  P.reserved2 = 0;
  ppk_mhd_params_init(&P);
  if (p) *p = P;

  if (pr) {
    ppk_mhd_problem R{};
    std::string prob = m.get_string("hydro", "problem", "unknown");
    if (prob == "blast") R.kind = PPK_PROBLEM_BLAST;
    else if (prob == "implode") R.kind = PPK_PROBLEM_IMPLODE;
    else if (prob == "kelvin_helmholtz") R.kind = PPK_PROBLEM_KELVIN_HELMHOLTZ;
    else if (prob == "wave") R.kind = PPK_PROBLEM_WAVE;
    // MHD only (SolverMHDMuscl.h:686-695); the hydro solver falls back to implode for them
    else if (prob == "rotor") R.kind = P.ghost_width == 2 ? PPK_PROBLEM_IMPLODE : PPK_PROBLEM_ROTOR;
    else if (prob == "field_loop" || prob == "field loop")
      R.kind = P.ghost_width == 2 ? PPK_PROBLEM_IMPLODE : PPK_PROBLEM_FIELD_LOOP;
    // hydro problems (SolverHydroMuscl.cpp:106-256): Gresho vortex (2D, 3D), isentropic
    // vortex and four quadrant (2D; the 3D solver falls back to implode for both); the gravity
    // problems, which this library does not provide, get kind -1, which
    // ppk_hydro_init_problem rejects
    else if (P.ghost_width == 2 && prob == "gresho_vortex") R.kind = PPK_PROBLEM_GRESHO_VORTEX;
    else if (P.ghost_width == 2 && prob == "isentropic_vortex")
      R.kind = P.dim == 2 ? PPK_PROBLEM_ISENTROPIC_VORTEX : PPK_PROBLEM_IMPLODE;
    else if (P.ghost_width == 2 && prob == "four_quadrant")
      R.kind = P.dim == 2 ? PPK_PROBLEM_FOUR_QUADRANT : PPK_PROBLEM_IMPLODE;
    else if (P.ghost_width == 2 && prob == "rayleigh_taylor") R.kind = PPK_PROBLEM_RAYLEIGH_TAYLOR;
    else if (P.ghost_width == 2 && prob == "rising_bubble") R.kind = PPK_PROBLEM_RISING_BUBBLE;
    else if (P.ghost_width == 2 && prob == "disk") R.kind = PPK_PROBLEM_DISK;
    // unknown names: the MHD solver falls back to Orszag-Tang (SolverMHDMuscl.h:701-709),
    // the hydro solver to implode (SolverHydroMuscl.cpp:170-177)
    else R.kind = P.ghost_width == 2 ? PPK_PROBLEM_IMPLODE : PPK_PROBLEM_ORSZAG_TANG;
    // OrszagTangParams
    R.ot_kt = m.get_float("OrszagTang", "kt", 0.0f);
    R.ot_b0_scale = m.get_float("OrszagTang", "b0_scale", 1.0f);
    // BlastParams (defaults computed in double, passed as float)
    double xmin = m.get_float("mesh", "xmin", 0.0f), ymin = m.get_float("mesh", "ymin", 0.0f);
    double zmin = m.get_float("mesh", "zmin", 0.0f), xmax = m.get_float("mesh", "xmax", 1.0f);
    double ymax = m.get_float("mesh", "ymax", 1.0f), zmax = m.get_float("mesh", "zmax", 1.0f);
    R.blast_radius = m.get_float("blast", "radius", (float)((xmin + xmax) / 2.0 / 10));
    R.blast_center_x = m.get_float("blast", "center_x", (float)((xmin + xmax) / 2));
    R.blast_center_y = m.get_float("blast", "center_y", (float)((ymin + ymax) / 2));
    R.blast_center_z = m.get_float("blast", "center_z", (float)((zmin + zmax) / 2));
    R.blast_density_in = m.get_float("blast", "density_in", 1.0f);
    R.blast_density_out = m.get_float("blast", "density_out", 1.2f);
    R.blast_pressure_in = m.get_float("blast", "pressure_in", 10.0f);
    R.blast_pressure_out = m.get_float("blast", "pressure_out", 0.1f);
    // uniform field: 0.5 in the reference, overridable (extension)
    R.blast_bx = m.get_string("blast", "bx", "").empty() ? 0.5 : m.get_float("blast", "bx", 0.5f);
    R.blast_by = m.get_string("blast", "by", "").empty() ? 0.5 : m.get_float("blast", "by", 0.5f);
    R.blast_bz = m.get_string("blast", "bz", "").empty() ? 0.5 : m.get_float("blast", "bz", 0.5f);
    // KHParams
    R.kh_d_in = m.get_float("KH", "d_in", 1.0f);
    R.kh_d_out = m.get_float("KH", "d_out", 1.0f);
    R.kh_pressure = m.get_float("KH", "pressure", 10.0f);
    R.kh_p_sine = m.get_bool("KH", "perturbation_sine", false);
    R.kh_p_sine_rob = m.get_bool("KH", "perturbation_sine_robertson", true);
    R.kh_p_rand = m.get_bool("KH", "perturbation_rand", false);
    R.kh_vflow_in = m.get_float("KH", "vflow_in", -0.5f);
    R.kh_vflow_out = m.get_float("KH", "vflow_out", 0.5f);
    R.kh_seed = (int)m.get_integer("KH", "rand_seed", 12);
    R.kh_amplitude = m.get_float("KH", "amplitude", 0.1f);
    R.kh_mode = (int)m.get_integer("KH", "mode", 2);
    R.kh_w0 = m.get_float("KH", "w0", 0.1f);
    R.kh_delta = m.get_float("KH", "delta", 0.03f);
    R.kh_axis = (int)m.get_integer("KH", "axis", 2);
    R.kh_bx = m.get_float("KH", "bx", 0.0f);
    R.kh_by = m.get_float("KH", "by", 0.0f);
    R.kh_bz = m.get_float("KH", "bz", 0.0f);
    // WaveParams re-reads mesh extents and gamma0 with its own defaults
    double wxmin = m.get_float("mesh", "xmin", 0.0f), wymin = m.get_float("mesh", "ymin", 0.0f);
    double wzmin = m.get_float("mesh", "zmin", 0.0f), wxmax = m.get_float("mesh", "xmax", 3.0f);
    double wymax = m.get_float("mesh", "ymax", 1.5f), wzmax = m.get_float("mesh", "zmax", 1.5f);
    R.wave_lx = wxmax - wxmin;
    R.wave_ly = wymax - wymin;
    R.wave_lz = wzmax - wzmin;
    R.wave_gamma0 = m.get_float("hydro", "gamma0", 1.66667f);
    R.wave_amplitude = m.get_float("wave", "amplitude", 1.0e-6f);
    R.wave_type = (int)m.get_integer("wave", "type", 0);
    // ImplodeParams
    const char *onames[8] = {"density_outer", "pressure_outer", "vx_outer", "vy_outer",
                             "vz_outer", "Bx_outer", "By_outer", "Bz_outer"};
    const char *inames[8] = {"density_inner", "pressure_inner", "vx_inner", "vy_inner",
                             "vz_inner", "Bx_inner", "By_inner", "Bz_inner"};
    const float odef[8] = {1.0f, 1.0f, 0, 0, 0, 0, 0, 0};
    const float idef[8] = {0.125f, 0.14f, 0, 0, 0, 0, 0, 0};
    for (int v = 0; v < 8; ++v) {
      R.implode_out[v] = m.get_float("implode", onames[v], odef[v]);
      R.implode_in[v] = m.get_float("implode", inames[v], idef[v]);
    }
    R.implode_shape = (int)m.get_integer("implode", "shape_region", 0);
    // GreshoParams (src/shared/problems/GreshoParams.h)
    R.gresho_rho0 = m.get_float("Gresho", "rho0", 1.0f);
    R.gresho_ma = m.get_float("Gresho", "Ma", 0.1f);
    // IsentropicVortexParams (src/shared/problems/IsentropicVortexParams.h): centre defaults
    // from the mesh extents, computed in double and passed as float defaults
    R.iv_rho_a = m.get_float("isentropic_vortex", "density_ambient", 1.0f);
    R.iv_p_a = m.get_float("isentropic_vortex", "pressure_ambient", 1.0f);
    R.iv_t_a = m.get_float("isentropic_vortex", "temperature_ambient", 1.0f);
    R.iv_u_a = m.get_float("isentropic_vortex", "vx_ambient", 1.0f);
    R.iv_v_a = m.get_float("isentropic_vortex", "vy_ambient", 1.0f);
    R.iv_w_a = m.get_float("isentropic_vortex", "vz_ambient", 1.0f);
    R.iv_x = m.get_float("isentropic_vortex", "center_x", (float)((xmin + xmax) / 2));
    R.iv_y = m.get_float("isentropic_vortex", "center_y", (float)((ymin + ymax) / 2));
    R.iv_z = m.get_float("isentropic_vortex", "center_z", (float)((zmin + zmax) / 2));
    R.iv_beta = m.get_float("isentropic_vortex", "strength", 5.0f);
    // RayleighTaylorInstabilityParams, RisingBubbleParams
    R.rti_amplitude = m.get_float("rayleigh_taylor", "amplitude", 0.01f);
    R.rti_d0 = m.get_float("rayleigh_taylor", "d0", 1.0f);
    R.rti_d1 = m.get_float("rayleigh_taylor", "d1", 2.0f);
    R.rti_gx = m.get_float("rayleigh_taylor", "gx", 0.0f);
    R.rti_gy = m.get_float("rayleigh_taylor", "gy", -0.1f);
    R.rti_gz = m.get_float("rayleigh_taylor", "gz", 0.0f);
    R.rb_R = m.get_float("rising_bubble", "R", 0.1f);
    R.rb_x0 = m.get_float("rising_bubble", "x0", 0.5f);
    R.rb_y0 = m.get_float("rising_bubble", "y0", 0.25f);
    R.rb_z0 = m.get_float("rising_bubble", "z0", 0.5f);
    R.rb_din = m.get_float("rising_bubble", "din", 1.0f);
    R.rb_dout = m.get_float("rising_bubble", "dout", 2.0f);
    R.rb_gx = m.get_float("rising_bubble", "gx", 0.0f);
    R.rb_gy = m.get_float("rising_bubble", "gy", -0.1f);
    R.rb_gz = m.get_float("rising_bubble", "gz", 0.0f);
    // DiskParams: its own mesh defaults (-0.5 .. 0.5), computed in double, passed as floats
    {
      const double dxmin = m.get_float("mesh", "xmin", -0.5f), dymin = m.get_float("mesh", "ymin", -0.5f);
      const double dzmin = m.get_float("mesh", "zmin", -0.5f), dxmax = m.get_float("mesh", "xmax", 0.5f);
      const double dymax = m.get_float("mesh", "ymax", 0.5f), dzmax = m.get_float("mesh", "zmax", 0.5f);
      R.disk_radius = m.get_float("disk", "radius", (float)((dxmax - dxmin) * 0.125));
      R.disk_radius_inner = m.get_float("disk", "radius_inner", (float)(R.disk_radius / 10));
      R.disk_xc = m.get_float("disk", "xc", (float)((dxmin + dxmax) / 2));
      R.disk_yc = m.get_float("disk", "yc", (float)((dymin + dymax) / 2));
      R.disk_zc = m.get_float("disk", "zc", (float)((dzmin + dzmax) / 2));
      R.disk_ref_density = m.get_float("disk", "ref_density", 1.0f);
      R.disk_contrast_density = m.get_float("disk", "contrast_density", 100.0f);
      R.disk_contrast_width = m.get_float("disk", "contrast_width", 0.01f);
      R.disk_ref_sound_speed = m.get_float("disk", "ref_sound_speed", 0.2f);
      // PointSourceGravity: GM = G * M of the two float values
      R.pg_x = m.get_float("gravity", "x", 0.0f);
      R.pg_y = m.get_float("gravity", "y", 0.0f);
      R.pg_z = m.get_float("gravity", "z", 0.0f);
      const double G = m.get_float("gravity", "G", 1.0f), M = m.get_float("gravity", "M", 1.0f);
      R.pg_gm = G * M;
      R.pg_eps = m.get_float("gravity", "soften", 0.01f);
    }
    // RotorParams (b0 default 5 / sqrt(4 pi) computed in double, passed as a float), FieldLoopParams
    R.rotor_r0 = m.get_float("rotor", "r0", 0.1f);
    R.rotor_r1 = m.get_float("rotor", "r1", 0.115f);
    R.rotor_u0 = m.get_float("rotor", "u0", 2.0f);
    R.rotor_p0 = m.get_float("rotor", "p0", 1.0f);
    R.rotor_b0 = m.get_float("rotor", "b0", (float)(5.0 / std::sqrt(4 * M_PI)));
    R.loop_radius = m.get_float("FieldLoop", "radius", 1.0f);
    R.loop_density_in = m.get_float("FieldLoop", "density_in", 1.0f);
    R.loop_amplitude = m.get_float("FieldLoop", "amplitude", 1.0f);
    R.loop_vflow = m.get_float("FieldLoop", "vflow", 1.0f);
    // four quadrant (SolverHydroMuscl.cpp:75-77)
    R.fq_config = (int)m.get_integer("riemann2d", "config_number", 0);
    R.fq_x = m.get_float("riemann2d", "x", 0.8f);
    R.fq_y = m.get_float("riemann2d", "y", 0.8f);
    *pr = R;
  }
  if (P.dim == 0) {
    set_thread_error("solver_name '" + sname +
                     "' is not handled by this library (MHD_Muscl_2D/3D, Hydro_Muscl_2D/3D)");
    return PPK_ERR_CONFIG;
  }
  return PPK_OK;
}

}  // namespace
}  // namespace ppk

extern "C" {

int ppk_mhd_params_init(ppk_mhd_params *p) {
  if (!p) return PPK_ERR_INVALID_ARGUMENT;
  // HydroParams::init, src/shared/HydroParams.cpp:421-453
  p->dx = (p->xmax - p->xmin) / p->nx;
  p->dy = (p->ymax - p->ymin) / p->ny;
  p->dz = (p->zmax - p->zmin) / p->nz;
  p->smallp = p->smallc * p->smallc / p->gamma0;
  p->smallpp = p->smallr * p->smallp;
  p->gamma6 = (p->gamma0 + 1.0) / (2.0 * p->gamma0);
  return PPK_OK;
}

int ppk_mhd_params_from_ini_text(const char *ini_text, ppk_mhd_params *params,
                                 ppk_mhd_problem *problem, char *solver_name, int name_len) {
  if (!ini_text) {
    ppk::set_thread_error("ini_text is NULL");
    return PPK_ERR_INVALID_ARGUMENT;
  }
  ppk::IniText m{std::string(ini_text)};
  return ppk::fill_from_ini(m, params, problem, solver_name, name_len);
}

int ppk_mhd_params_from_ini(const char *ini_path, ppk_mhd_params *params,
                            ppk_mhd_problem *problem, char *solver_name, int name_len) {
  if (!ini_path) {
    ppk::set_thread_error("ini_path is NULL");
    return PPK_ERR_INVALID_ARGUMENT;
  }
  std::ifstream f(ini_path);
  if (!f) {
    ppk::set_thread_error(std::string("cannot open parameter file ") + ini_path);
    return PPK_ERR_CONFIG;
  }
  std::stringstream ss;
  ss << f.rdbuf();
  ppk::IniText m{ss.str()};
  return ppk::fill_from_ini(m, params, problem, solver_name, name_len);
}

int ppk_mhd_params_from_lookup(ppk_ini_lookup lookup, void *ctx, ppk_mhd_params *params,
                               ppk_mhd_problem *problem, char *solver_name, int name_len) {
  if (!lookup) {
    ppk::set_thread_error("lookup is NULL");
    return PPK_ERR_INVALID_ARGUMENT;
  }
  ppk::IniCallback m{lookup, ctx};
  return ppk::fill_from_ini(m, params, problem, solver_name, name_len);
}

}  // extern "C"
