// C ABI of the library: solver handles, state transfers and the time loop
// (include/ppkmhd_amd.h).  Host orchestration only; kernels live in the .hip files.
//
// Time-loop state lives on the device (ctrl block: t, dt, CFL accumulator, iteration), so a
// run of many steps is launched without any host synchronization; the host reads the block
// back only when asked (get_time, finished, compute_dt).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "capi_internal.hpp"
#include "event_timer.hpp"
#include "kernels.hpp"
#include "ppkmhd_amd.h"

namespace ppk {

namespace {
thread_local std::string g_thread_error;

}  // namespace
const KernelTable kFast = {fast::staged_boundaries, fast::staged_prim_dt,
                           fast::staged_begin,      fast::staged_set_dt,
                           fast::staged_end,        fast::staged_godunov,
                           fast::fused_godunov,     fast::fused_cfl,
                           fast::fused_half,        fast::surrogate_dump,
                           fast::surrogate_dump_slots, fast::march_godunov,
                           fast::batch_block_counts,
                           fast::batch_step};
const KernelTable kStrict = {strict::staged_boundaries, strict::staged_prim_dt,
                             strict::staged_begin,      strict::staged_set_dt,
                             strict::staged_end,        strict::staged_godunov,
                             strict::fused_godunov,     strict::fused_cfl,
                             strict::fused_half,        strict::surrogate_dump,
                             strict::surrogate_dump_slots, strict::march_godunov,
                             strict::batch_block_counts,
                             strict::batch_step};

void set_thread_error(const std::string &msg) { g_thread_error = msg; }
const char *thread_error() { return g_thread_error.c_str(); }

// diagnostics (diag.hip)
hipError_t diag_divb(const DevParams &p, const double *U, double *out, hipStream_t s);
hipError_t diag_math(const double *x, long long n, double *out, hipStream_t s);
hipError_t diag_sums(const DevParams &p, const double *U, double *partials, int nblocks,
                     hipStream_t s);
constexpr int kDiagBlocks = 512;

const char *const kStageNames[ST_COUNT] = {
    "boundaries", "primitives_dt", "elec_field", "mag_slopes",    "trace",        "fluxes", "emf",
    "update_ct",  "control",       "fused_half", "fused_riemann", "fused_update", "fused_cfl",
    "fused_march", "surrogate_faces", "surrogate_fallback"};
static_assert(ST_COUNT == 16, "stage names");

// the fused kernels serve every precision in 2D and 3D unless the staged path is asked for
bool fused_supported(const DevParams &p, int precision, int kernel_path) {
  (void)p;
  if (kernel_path == PPK_PATH_STAGED) return false;
  return precision == PPK_FP64 || precision == PPK_FP64_STRICT || precision == PPK_FP32_BULK;
}

}  // namespace ppk

struct ppk_mhd_solver {
  ppk_mhd_params p{};
  ppk::DevParams dp{};
  hipStream_t stream = nullptr;
  int device = 0;
  double *U[2] = {nullptr, nullptr};
  double *Q = nullptr;
  double *ctrl = nullptr;       // device control block
  double *ctrl_host = nullptr;  // pinned mirror (CTRL_SLOTS) + one diagnostics slot
  double *diag = nullptr;       // diagnostics scratch (device)
  double *parts = nullptr;      // per-block CFL partial maxima (device)
  ppk::StagedScratch w;
  ppk::FusedScratch fw;
  ppk::SurrogateCfg sur;                   // surrogate face solves (fw.sur points here when on)
  unsigned long long *sur_counts = nullptr;  // device: faces by the network, by HLLD
  unsigned char *sur_flags = nullptr;        // device: faces left to the HLLD pass
  bool fused = false;
  // parts holds the CFL maxima of cur() (ghosts as the next fill makes them): set by a fused
  // CFL-mode step or compute_dt, cleared by anything that changes cur() behind the kernels
  bool cfl_valid = false;
  int nparts = 0;  // how many partial maxima the kernel that last wrote parts produced
  const ppk::KernelTable *K = nullptr;
  ppk::EventTimer *timer = nullptr;
  long long iter_host = 0;  // iteration count as far as the host knows (buffer parity)
  std::string err;

  ppk::StageTimer *tm() { return (timer && timer->mask()) ? timer : nullptr; }
  double *cur() { return U[iter_host & 1]; }
  double *nxt() { return U[(iter_host + 1) & 1]; }
  long long cells() const { return (long long)dp.isize * dp.jsize * dp.ksize; }
  size_t state_bytes() const { return (size_t)cells() * 8 * sizeof(double); }
};

namespace {

int fail(ppk_mhd_t *h, int code, const std::string &msg) {
  if (h)
    h->err = msg;
  else
    ppk::set_thread_error(msg);
  return code;
}

#define HIP_CHECK(h, expr)                                                               \
  do {                                                                                   \
    hipError_t _e = (expr);                                                              \
    if (_e != hipSuccess)                                                                \
      return fail((h), PPK_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e));  \
  } while (0)

int validate(const ppk_mhd_params *p, std::string &why) {
  if (!p) { why = "params is NULL"; return PPK_ERR_INVALID_ARGUMENT; }
  if (p->dim != 2 && p->dim != 3) { why = "dim must be 2 or 3"; return PPK_ERR_INVALID_ARGUMENT; }
  if (p->ghost_width != 3) {
    why = "ghost_width must be 3 for MHD (HydroParams.cpp:56-71)";
    return PPK_ERR_INVALID_ARGUMENT;
  }
  if (p->implementation_version < 0 || p->implementation_version > 2) {
    why = "implementationVersion must be 0, 1 or 2 (HydroParams.cpp:446-451)";
    return PPK_ERR_INVALID_ARGUMENT;
  }
  if (p->implementation_version == 2) {
    why = "implementationVersion 2 (the updated-primitive scheme of SolverMHDMuscl.cpp:430-456) "
          "is not provided; versions 0 and 1 run the version-0 scheme";
    return PPK_ERR_UNSUPPORTED;
  }
  if (p->nx < 3 || p->ny < 3 || (p->dim == 3 && p->nz < 3)) {
    why = "nx, ny (and nz in 3D) must be >= ghost_width";
    return PPK_ERR_INVALID_ARGUMENT;
  }
  if (p->dim == 2 && p->nz != 1) { why = "nz must be 1 in 2D"; return PPK_ERR_INVALID_ARGUMENT; }
  long long cells = (long long)(p->nx + 6) * (p->ny + 6) * (p->dim == 3 ? p->nz + 6 : 1);
  if (cells > (1LL << 31)) { why = "grid too large (more than 2^31 padded cells)"; return PPK_ERR_INVALID_ARGUMENT; }
  if (p->riemann != PPK_RIEMANN_HLLD && p->riemann != PPK_RIEMANN_HLL &&
      p->riemann != PPK_RIEMANN_LLF) {
    why = "riemann must be hlld, hll or llf for MHD (RiemannSolvers_MHD.h:372-392)";
    return PPK_ERR_UNSUPPORTED;
  }
  if (p->slope_type != 1.0 && p->slope_type != 2.0) {
    why = "slope_type must be 1 (minmod) or 2 (MC)";
    return PPK_ERR_UNSUPPORTED;
  }
  int nf = p->dim == 3 ? 6 : 4;
  for (int f = 0; f < nf; ++f)
    if (p->bc[f] != PPK_BC_DIRICHLET && p->bc[f] != PPK_BC_NEUMANN && p->bc[f] != PPK_BC_PERIODIC) {
      why = "boundary types must be 1 (Dirichlet), 2 (Neumann) or 3 (periodic)";
      return PPK_ERR_INVALID_ARGUMENT;
    }
  if (p->precision != PPK_FP64 && p->precision != PPK_FP64_STRICT && p->precision != PPK_FP32_BULK) {
    why = "unknown precision";
    return PPK_ERR_INVALID_ARGUMENT;
  }
  if (p->kernel_path < PPK_PATH_AUTO || p->kernel_path > PPK_PATH_FUSED_RECORDS) {
    why = "unknown kernel_path";
    return PPK_ERR_INVALID_ARGUMENT;
  }
  if (!(p->gamma0 > 1.0) || !(p->dx > 0) || !(p->dy > 0) || (p->dim == 3 && !(p->dz > 0))) {
    why = "gamma0 must be > 1 and dx, dy, dz > 0 (call ppk_mhd_params_init)";
    return PPK_ERR_INVALID_ARGUMENT;
  }
  return PPK_OK;
}

ppk::DevParams make_dev(const ppk_mhd_params &p) {
  ppk::DevParams d{};
  d.dim = p.dim;
  d.nx = p.nx;
  d.ny = p.ny;
  d.nz = p.dim == 3 ? p.nz : 1;
  d.gw = p.ghost_width;
  d.isize = p.nx + 2 * d.gw;
  d.jsize = p.ny + 2 * d.gw;
  d.ksize = p.dim == 3 ? p.nz + 2 * d.gw : 1;
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
  // the single-grid fused step runs one configuration: the 16 x 4 x 4 half-step tile, runs of
  // 256 cells in 4 x 4 row patches for the per-cell kernels, all face and edge solves in one
  // launch (the batch selects its own tile and brick shapes, ppk_mhd_batch_autotune)
  d.tile3 = 0;
  d.brick = 0;
  return d;
}

int write_ctrl(ppk_mhd_t *h) {
  HIP_CHECK(h, hipMemcpyAsync(h->ctrl, h->ctrl_host, sizeof(double) * ppk::CTRL_SLOTS,
                              hipMemcpyHostToDevice, h->stream));
  HIP_CHECK(h, hipStreamSynchronize(h->stream));
  return PPK_OK;
}

int read_ctrl(ppk_mhd_t *h) {
  HIP_CHECK(h, hipMemcpyAsync(h->ctrl_host, h->ctrl, sizeof(double) * ppk::CTRL_SLOTS,
                              hipMemcpyDeviceToHost, h->stream));
  HIP_CHECK(h, hipStreamSynchronize(h->stream));
  return PPK_OK;
}

// device buffers of the three-kernel fused form (records, fluxes, EMFs)
bool alloc_records(ppk_mhd_t *h) {
  if (h->fw.rec) return true;
  const int d = h->dp.dim;
  const size_t n = (size_t)h->cells();
  const size_t es = h->dp.bulk32 ? sizeof(float) : sizeof(double);
  return hipMalloc(&h->fw.rec, (size_t)(d == 3 ? ppk::kRec3 : ppk::kRec2) * n * es) == hipSuccess &&
         hipMalloc(&h->fw.flux,
                   (size_t)d * (d == 3 ? ppk::kFluxComp3 : ppk::kFluxComp2) * n * es) == hipSuccess &&
         hipMalloc(&h->fw.emf, (size_t)(d == 3 ? 3 : 1) * n * sizeof(double)) == hipSuccess;
}

// one time step from cur() into nxt() (godunov_unsplit_impl), optionally gated on finished()
int launch_step(ppk_mhd_t *h, double dt, int dt_mode, int check_finished) {
  const ppk::KernelTable &K = *h->K;
  const ppk::DevParams &p = h->dp;
  hipStream_t s = h->stream;
  double *in = h->cur(), *out = h->nxt();
  ppk::StageTimer *tm = h->tm();
  const bool cfl = dt_mode == PPK_DT_CFL;
  const int mode = cfl ? ppk::SETDT_CFL : ppk::SETDT_FIXED;
  HIP_CHECK(h, K.begin(h->ctrl, check_finished, s));
  // the single-pass step writes only the interior of its output: the ghost fill gives the
  // output the filled input's ghosts (the reference's U2 = U copy) in the same pass
  const bool march = h->fused && p.march > 0 && !h->fw.sur;
  HIP_CHECK(h, K.boundaries(p, in, march ? out : nullptr, h->ctrl, s, tm));
  if (h->fused) {
    // the previous fused step left the CFL maxima of its output (= in) in parts
    if (cfl && !h->cfl_valid) {
      HIP_CHECK(h, K.fused_cfl(p, in, h->parts, h->ctrl, s, tm));
      h->nparts = ppk::fused_cfl_parts(p);
    }
    HIP_CHECK(h, K.set_dt(h->ctrl, h->parts, h->nparts, dt, mode, h->p.cfl, s));
    if (march) {  // the single-pass 3D step
      HIP_CHECK(h, K.march_godunov(p, in, out, h->parts, h->ctrl, cfl ? 1 : 0, s, tm));
      if (cfl) h->nparts = (int)ppk::march_blocks(p);
    } else {
      HIP_CHECK(h, K.fused_godunov(p, in, out, h->fw, h->parts, h->ctrl, cfl ? 1 : 0, s, tm));
      if (cfl) h->nparts = ppk::fused_cfl_parts(p);
    }
    h->cfl_valid = cfl;
  } else {
    const int nparts = ppk::cfl_blocks(p);
    HIP_CHECK(h, K.prim_dt(p, in, h->Q, h->parts, h->ctrl, cfl ? 1 : 0, s, tm));
    HIP_CHECK(h, K.set_dt(h->ctrl, h->parts, nparts, dt, mode, h->p.cfl, s));
    HIP_CHECK(h, K.godunov(p, in, out, h->Q, h->w, h->ctrl, s, tm));
  }
  HIP_CHECK(h, K.end(h->ctrl, s));
  return PPK_OK;
}

void free_all(ppk_mhd_t *h) {
  auto F = [](void *ptr) {
    if (ptr) (void)hipFree(ptr);
  };
  F(h->U[0]);
  F(h->U[1]);
  F(h->Q);
  F(h->ctrl);
  F(h->diag);
  F(h->parts);
  F(h->w.faces);
  F(h->w.edges);
  F(h->w.flux);
  F(h->w.emf);
  F(h->w.elec);
  F(h->w.dbf);
  F(h->fw.rec);
  F(h->fw.flux);
  F(h->fw.emf);
  F(h->sur_counts);
  F(h->sur_flags);
  delete h->timer;
  h->timer = nullptr;
  if (h->ctrl_host) (void)hipHostFree(h->ctrl_host);
  if (h->stream) (void)hipStreamDestroy(h->stream);
}

}  // namespace

namespace ppk {
int validate_params(const ppk_mhd_params *p, std::string &why) { return validate(p, why); }
// the MHD solvers' problems (SolverMHDMuscl.h:672-709): the hydro-only kinds are refused
bool mhd_problem_ok(int kind) {
  return (kind >= PPK_PROBLEM_ORSZAG_TANG && kind <= PPK_PROBLEM_IMPLODE) ||
         kind == PPK_PROBLEM_ROTOR || kind == PPK_PROBLEM_FIELD_LOOP;
}
DevParams make_dev_params(const ppk_mhd_params &p) { return make_dev(p); }
ppk_mhd_params handle_params(const ppk_mhd_t *h) { return h->p; }
void set_handle_error(ppk_mhd_t *h, const std::string &msg) { h->err = msg; }
}  // namespace ppk

extern "C" {

int ppk_mhd_params_size(void) { return (int)sizeof(ppk_mhd_params); }
int ppk_mhd_problem_size(void) { return (int)sizeof(ppk_mhd_problem); }
const char *ppk_mhd_version(void) { return "ppkmhd_amd 0.1 (gfx950)"; }

int ppk_mhd_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

const char *ppk_mhd_last_error(const ppk_mhd_t *h) {
  return h ? h->err.c_str() : ppk::thread_error();
}

int ppk_mhd_create(const ppk_mhd_params *params, ppk_mhd_t **handle) {
  if (!handle) return fail(nullptr, PPK_ERR_INVALID_ARGUMENT, "handle is NULL");
  *handle = nullptr;
  std::string why;
  int rc = validate(params, why);
  if (rc != PPK_OK) return fail(nullptr, rc, why);
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return fail(nullptr, PPK_ERR_HIP, "no HIP device available");

  ppk_mhd_t *h = new (std::nothrow) ppk_mhd_t();
  if (!h) return fail(nullptr, PPK_ERR_OUT_OF_MEMORY, "host allocation failed");
  h->p = *params;
  h->dp = make_dev(*params);
  h->K = params->precision == PPK_FP64_STRICT ? &ppk::kStrict : &ppk::kFast;
  h->dp.bulk32 = params->precision == PPK_FP32_BULK ? 1 : 0;
  h->fused = ppk::fused_supported(h->dp, params->precision, params->kernel_path);
  // FP64 3D: the single-pass kernel, unless the three-kernel form is asked for (the FP32-bulk
  // form and the surrogate run the three-kernel form)
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    cus = 0;
  h->dp.march = (params->dim == 3 && h->fused && params->precision != PPK_FP32_BULK &&
                 params->kernel_path != PPK_PATH_FUSED_RECORDS)
                    ? ppk::march_nblk(h->dp, cus)
                    : 0;
  if (params->kernel_path == PPK_PATH_FUSED && !h->fused) {
    delete h;
    return fail(nullptr, PPK_ERR_UNSUPPORTED, "fused kernels do not support these parameters");
  }
  if (params->precision == PPK_FP32_BULK && !h->fused) {
    delete h;
    return fail(nullptr, PPK_ERR_UNSUPPORTED, "FP32 bulk arithmetic needs the fused kernels");
  }
  auto bail = [&](int code, const std::string &m) {
    free_all(h);
    delete h;
    return fail(nullptr, code, m);
  };
  if (hipGetDevice(&h->device) != hipSuccess) return bail(PPK_ERR_HIP, "hipGetDevice failed");
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess)
    return bail(PPK_ERR_HIP, "hipStreamCreateWithFlags failed");
  h->timer = new (std::nothrow) ppk::EventTimer(h->stream);
  if (!h->timer) return bail(PPK_ERR_OUT_OF_MEMORY, "host allocation failed");

  const size_t n = (size_t)h->cells();
  auto dalloc = [&](double **ptr, size_t count) {
    return hipMalloc((void **)ptr, count * sizeof(double)) == hipSuccess;
  };
  bool ok = dalloc(&h->U[0], 8 * n) && dalloc(&h->U[1], 8 * n) &&
            dalloc(&h->ctrl, ppk::CTRL_SLOTS) && dalloc(&h->diag, ppk::kDiagBlocks * 8 + 8) &&
            dalloc(&h->parts, std::max<size_t>({(size_t)ppk::kCflMaxBlocks,
                                                (size_t)ppk::fused_cfl_parts(h->dp),
                                                (size_t)ppk::march_blocks(h->dp)}));
  if (ok && !h->fused) {
    const int d = params->dim;
    ok = dalloc(&h->Q, 8 * n) && dalloc(&h->w.faces, (size_t)2 * d * 8 * n) &&
         dalloc(&h->w.edges, (size_t)4 * (d == 3 ? 3 : 1) * 8 * n) &&
         dalloc(&h->w.flux, (size_t)d * 8 * n) && dalloc(&h->w.emf, (size_t)(d == 3 ? 3 : 1) * n);
    if (ok && d == 3) ok = dalloc(&h->w.elec, 3 * n) && dalloc(&h->w.dbf, 6 * n);
  }
  // the three-kernel form's records, fluxes and EMFs; the single-pass 3D step needs none of
  // them (they are allocated when the surrogate or a face dump first asks for them)
  if (ok && h->fused && h->dp.march == 0) ok = alloc_records(h);
  if (!ok) return bail(PPK_ERR_OUT_OF_MEMORY, "device allocation failed");
  if (hipHostMalloc((void **)&h->ctrl_host, sizeof(double) * (ppk::CTRL_SLOTS + 1)) != hipSuccess)
    return bail(PPK_ERR_OUT_OF_MEMORY, "pinned host allocation failed");
  if (hipMemsetAsync(h->U[0], 0, 8 * n * sizeof(double), h->stream) != hipSuccess ||
      hipMemsetAsync(h->U[1], 0, 8 * n * sizeof(double), h->stream) != hipSuccess)
    return bail(PPK_ERR_HIP, "hipMemsetAsync failed");
  std::fill(h->ctrl_host, h->ctrl_host + ppk::CTRL_SLOTS, 0.0);
  // SolverBase::read_config: t = tCurrent (0), dt = tEnd (src/shared/SolverBase.cpp:119-121)
  h->ctrl_host[ppk::CTRL_T] = 0.0;
  h->ctrl_host[ppk::CTRL_DT] = params->tEnd;
  h->ctrl_host[ppk::CTRL_ITER] = 0.0;
  h->ctrl_host[ppk::CTRL_TEND] = params->tEnd;
  h->ctrl_host[ppk::CTRL_NSTEPMAX] = (double)params->n_step_max;
  h->ctrl_host[ppk::CTRL_ACTIVE] = 1.0;
  if (write_ctrl(h) != PPK_OK) {
    std::string m = h->err;
    return bail(PPK_ERR_HIP, m);
  }
  *handle = h;
  return PPK_OK;
}

void ppk_mhd_destroy(ppk_mhd_t *h) {
  if (!h) return;
  (void)hipStreamSynchronize(h->stream);
  free_all(h);
  delete h;
}

namespace {
// after the initial state is in U[0]: the rest of the SolverMHDMuscl constructor
// (boundaries, U2 = U, primitives, compute_dt; src/muscl/SolverMHDMuscl.h:389-402)
int finish_init(ppk_mhd_t *h);
}  // namespace

int ppk_mhd_init_problem(ppk_mhd_t *h, const ppk_mhd_problem *problem) {
  if (!h || !problem) return fail(h, PPK_ERR_INVALID_ARGUMENT, "NULL argument");
  if (!ppk::mhd_problem_ok(problem->kind))
    return fail(h, PPK_ERR_INVALID_ARGUMENT, "unknown problem kind");
  std::vector<double> host((size_t)h->cells() * 8);
  ppk::init_problem_host(h->p, *problem, host.data());
  h->iter_host = 0;
  h->cfl_valid = false;
  HIP_CHECK(h, hipMemcpyAsync(h->cur(), host.data(), h->state_bytes(), hipMemcpyHostToDevice,
                              h->stream));
  // the host array must stay alive until the copy has read it
  HIP_CHECK(h, hipStreamSynchronize(h->stream));
  return finish_init(h);
}

int ppk_mhd_init_problem_device(ppk_mhd_t *h, const ppk_mhd_problem *problem) {
  if (!h || !problem) return fail(h, PPK_ERR_INVALID_ARGUMENT, "NULL argument");
  if (!ppk::mhd_problem_ok(problem->kind))
    return fail(h, PPK_ERR_INVALID_ARGUMENT, "unknown problem kind");
  h->iter_host = 0;
  h->cfl_valid = false;
  // the other state buffer is the wave problem's vector-potential scratch (3 of its 8 planes)
  HIP_CHECK(h, ppk::init_problem_device(h->p, *problem, h->cur(), h->nxt(), h->stream));
  return finish_init(h);
}

int ppk_mhd_initial_state(const ppk_mhd_params *params, const ppk_mhd_problem *problem,
                          double *U, size_t count) {
  std::string why;
  int rc = ppk::validate_params(params, why);
  if (rc != PPK_OK) return fail(nullptr, rc, why);
  if (!problem || !U) return fail(nullptr, PPK_ERR_INVALID_ARGUMENT, "NULL argument");
  if (!ppk::mhd_problem_ok(problem->kind))
    return fail(nullptr, PPK_ERR_INVALID_ARGUMENT, "unknown problem kind");
  const size_t cells = (size_t)(params->nx + 2 * params->ghost_width) *
                       (params->ny + 2 * params->ghost_width) *
                       (params->dim == 3 ? params->nz + 2 * params->ghost_width : 1);
  if (count != 8 * cells)
    return fail(nullptr, PPK_ERR_INVALID_ARGUMENT, "count must be 8*isize*jsize*ksize");
  ppk::init_problem_host(*params, *problem, U);
  return PPK_OK;
}

namespace {
int finish_init(ppk_mhd_t *h) {
  h->ctrl_host[ppk::CTRL_T] = 0.0;
  h->ctrl_host[ppk::CTRL_ITER] = 0.0;
  h->ctrl_host[ppk::CTRL_ACTIVE] = 1.0;
  int rc = write_ctrl(h);
  if (rc != PPK_OK) return rc;
  double dt = 0.0;
  rc = ppk_mhd_compute_dt(h, &dt);
  if (rc != PPK_OK) return rc;
  HIP_CHECK(h, hipMemcpyAsync(h->nxt(), h->cur(), h->state_bytes(), hipMemcpyDeviceToDevice,
                              h->stream));
  h->ctrl_host[ppk::CTRL_DT] = (0.0 + dt > h->p.tEnd) ? h->p.tEnd - 0.0 : dt;
  return write_ctrl(h);
}
}  // namespace

int ppk_mhd_upload(ppk_mhd_t *h, const double *U, size_t count) {
  if (!h || !U) return fail(h, PPK_ERR_INVALID_ARGUMENT, "NULL argument");
  if (count != (size_t)h->cells() * 8)
    return fail(h, PPK_ERR_INVALID_ARGUMENT, "count must be 8*isize*jsize*ksize");
  h->cfl_valid = false;
  HIP_CHECK(h, hipMemcpyAsync(h->cur(), U, h->state_bytes(), hipMemcpyHostToDevice, h->stream));
  HIP_CHECK(h, hipStreamSynchronize(h->stream));
  return PPK_OK;
}

int ppk_mhd_download(ppk_mhd_t *h, double *U, size_t count) {
  if (!h || !U) return fail(h, PPK_ERR_INVALID_ARGUMENT, "NULL argument");
  if (count != (size_t)h->cells() * 8)
    return fail(h, PPK_ERR_INVALID_ARGUMENT, "count must be 8*isize*jsize*ksize");
  HIP_CHECK(h, hipMemcpyAsync(U, h->cur(), h->state_bytes(), hipMemcpyDeviceToHost, h->stream));
  HIP_CHECK(h, hipStreamSynchronize(h->stream));
  return PPK_OK;
}

int ppk_mhd_device_state(ppk_mhd_t *h, double **dU) {
  if (!h || !dU) return fail(h, PPK_ERR_INVALID_ARGUMENT, "NULL argument");
  h->cfl_valid = false;  // the caller may write the state
  *dU = h->cur();
  return PPK_OK;
}

int ppk_mhd_stream(ppk_mhd_t *h, void **stream) {
  if (!h || !stream) return fail(h, PPK_ERR_INVALID_ARGUMENT, "NULL argument");
  *stream = (void *)h->stream;
  return PPK_OK;
}

int ppk_mhd_make_boundaries(ppk_mhd_t *h) {
  if (!h) return fail(h, PPK_ERR_INVALID_ARGUMENT, "NULL handle");
  HIP_CHECK(h, h->K->begin(h->ctrl, 0, h->stream));
  HIP_CHECK(h, h->K->boundaries(h->dp, h->cur(), nullptr, h->ctrl, h->stream, h->tm()));
  return PPK_OK;
}

int ppk_mhd_compute_dt(ppk_mhd_t *h, double *dt) {
  if (!h || !dt) return fail(h, PPK_ERR_INVALID_ARGUMENT, "NULL argument");
  // the dt the next step would take: ghost fill, primitives, CFL max (no tEnd clamp)
  HIP_CHECK(h, h->K->begin(h->ctrl, 0, h->stream));
  HIP_CHECK(h, h->K->boundaries(h->dp, h->cur(), nullptr, h->ctrl, h->stream, h->tm()));
  if (h->fused) {
    HIP_CHECK(h, h->K->fused_cfl(h->dp, h->cur(), h->parts, h->ctrl, h->stream, h->tm()));
    h->cfl_valid = true;
  } else {
    HIP_CHECK(h, h->K->prim_dt(h->dp, h->cur(), h->Q, h->parts, h->ctrl, 1, h->stream, h->tm()));
  }
  const int nparts = h->fused ? ppk::fused_cfl_parts(h->dp) : ppk::cfl_blocks(h->dp);
  if (h->fused) h->nparts = nparts;
  HIP_CHECK(h, h->K->set_dt(h->ctrl, h->parts, nparts, 0.0,
                            ppk::SETDT_REDUCE_ONLY, h->p.cfl, h->stream));
  int rc = read_ctrl(h);
  if (rc != PPK_OK) return rc;
  *dt = h->p.cfl / h->ctrl_host[ppk::CTRL_INVDT];
  return PPK_OK;
}

int ppk_mhd_dump_faces(ppk_mhd_t *h, ppk_faceset_t *set, int stride, int phase,
                       size_t *appended) {
  if (!h || !set) return fail(h, PPK_ERR_INVALID_ARGUMENT, "NULL argument");
  if (!h->fused) return fail(h, PPK_ERR_UNSUPPORTED, "face dumps need the fused kernel path");
  if (stride < 1 || phase < 0 || phase >= stride)
    return fail(h, PPK_ERR_INVALID_ARGUMENT, "need stride >= 1 and 0 <= phase < stride");
  int finished = 0;
  int rc = ppk_mhd_finished(h, &finished);
  if (rc != PPK_OK) return rc;
  if (finished) return fail(h, PPK_ERR_INVALID_ARGUMENT, "the solver has finished");
  const long long n = h->K->surrogate_dump_slots(h->dp, stride, phase);
  if ((size_t)n > set->capacity - set->n)
    return fail(h, PPK_ERR_INVALID_ARGUMENT, "face set too small for this dump");
  HIP_CHECK(h, hipStreamSynchronize(h->stream));
  if (!alloc_records(h)) return fail(h, PPK_ERR_OUT_OF_MEMORY, "device allocation failed");
  // the faces the next CFL step will solve: ghost fill, CFL dt, half-step records
  HIP_CHECK(h, h->K->begin(h->ctrl, 1, h->stream));
  HIP_CHECK(h, h->K->boundaries(h->dp, h->cur(), nullptr, h->ctrl, h->stream, h->tm()));
  HIP_CHECK(h, h->K->fused_cfl(h->dp, h->cur(), h->parts, h->ctrl, h->stream, h->tm()));
  HIP_CHECK(h, h->K->set_dt(h->ctrl, h->parts, ppk::fused_cfl_parts(h->dp), 0.0,
                            ppk::SETDT_CFL, h->p.cfl, h->stream));
  HIP_CHECK(h, h->K->fused_half(h->dp, h->cur(), h->fw.rec, h->ctrl, h->stream));
  HIP_CHECK(h, h->K->surrogate_dump(h->dp, h->fw.rec, h->ctrl, set->X + set->n * ppk::kMlpIn,
                                    set->T + set->n * ppk::kMlpOut, stride, phase, h->stream));
  HIP_CHECK(h, hipStreamSynchronize(h->stream));
  h->cfl_valid = true;
  h->nparts = ppk::fused_cfl_parts(h->dp);
  set->n += (size_t)n;
  if (appended) *appended = (size_t)n;
  return PPK_OK;
}

int ppk_mhd_set_surrogate(ppk_mhd_t *h, const ppk_mlp_t *net, double gate_rho, double gate_p) {
  if (!h) return fail(h, PPK_ERR_INVALID_ARGUMENT, "NULL handle");
  // steps already queued keep the network (and gate) they were launched with: let them finish
  // before the configuration changes, so the caller may then free the previous network
  HIP_CHECK(h, hipStreamSynchronize(h->stream));
  if (!net) {
    h->fw.sur = nullptr;
    h->sur = ppk::SurrogateCfg{};
    return PPK_OK;
  }
  if (!h->fused) return fail(h, PPK_ERR_UNSUPPORTED, "the surrogate needs the fused kernel path");
  if (h->p.riemann != PPK_RIEMANN_HLLD)
    return fail(h, PPK_ERR_UNSUPPORTED, "the surrogate replaces HLLD: riemann must be hlld");
  if (!(gate_rho >= 0.0) || !(gate_p >= 0.0))
    return fail(h, PPK_ERR_INVALID_ARGUMENT, "gate thresholds must be >= 0");
  if (!alloc_records(h)) return fail(h, PPK_ERR_OUT_OF_MEMORY, "device allocation failed");
  if (!h->sur_counts) {
    const size_t bytes = 2 * ppk::kSurrogateBlocks * sizeof(unsigned long long);
    HIP_CHECK(h, hipMalloc(&h->sur_counts, bytes));
    HIP_CHECK(h, hipMemsetAsync(h->sur_counts, 0, bytes, h->stream));
  }
  if (!h->sur_flags)
    HIP_CHECK(h, hipMalloc(&h->sur_flags, (size_t)h->dp.dim * h->cells()));
  h->sur.flags = h->sur_flags;
  h->sur.net = net->packed;
  h->sur.gate_rho = gate_rho;
  h->sur.gate_p = gate_p;
  h->sur.counters = h->sur_counts;
  h->fw.sur = &h->sur;
  return PPK_OK;
}

int ppk_mhd_surrogate_counts(ppk_mhd_t *h, int64_t *by_network, int64_t *by_hlld, int reset) {
  if (!h) return fail(h, PPK_ERR_INVALID_ARGUMENT, "NULL handle");
  unsigned long long c[2] = {0, 0};
  if (h->sur_counts) {
    std::vector<unsigned long long> per(2 * ppk::kSurrogateBlocks);
    const size_t bytes = per.size() * sizeof(unsigned long long);
    HIP_CHECK(h, hipMemcpyAsync(per.data(), h->sur_counts, bytes, hipMemcpyDeviceToHost,
                                h->stream));
    HIP_CHECK(h, hipStreamSynchronize(h->stream));
    for (int b = 0; b < ppk::kSurrogateBlocks; ++b) {
      c[0] += per[2 * b];
      c[1] += per[2 * b + 1];
    }
    if (reset) {
      HIP_CHECK(h, hipMemsetAsync(h->sur_counts, 0, bytes, h->stream));
      HIP_CHECK(h, hipStreamSynchronize(h->stream));
    }
  }
  if (by_network) *by_network = (int64_t)c[0];
  if (by_hlld) *by_hlld = (int64_t)c[1];
  return PPK_OK;
}

int ppk_mhd_step(ppk_mhd_t *h, double dt, int dt_mode) {
  if (!h) return fail(h, PPK_ERR_INVALID_ARGUMENT, "NULL handle");
  if (dt_mode != PPK_DT_FIXED && dt_mode != PPK_DT_CFL)
    return fail(h, PPK_ERR_INVALID_ARGUMENT, "dt_mode must be PPK_DT_FIXED or PPK_DT_CFL");
  int rc = launch_step(h, dt, dt_mode, 0);
  if (rc != PPK_OK) return rc;
  h->iter_host += 1;
  return PPK_OK;
}

int ppk_mhd_run(ppk_mhd_t *h, int nsteps, double dt, int dt_mode) {
  if (!h) return fail(h, PPK_ERR_INVALID_ARGUMENT, "NULL handle");
  if (nsteps < 0) return fail(h, PPK_ERR_INVALID_ARGUMENT, "nsteps must be >= 0");
  if (dt_mode != PPK_DT_FIXED && dt_mode != PPK_DT_CFL)
    return fail(h, PPK_ERR_INVALID_ARGUMENT, "dt_mode must be PPK_DT_FIXED or PPK_DT_CFL");
  // steps after finished() are no-ops on the device; the buffer parity is re-read at the end
  const long long start = h->iter_host;
  const bool cfl_valid0 = h->cfl_valid;
  const int nparts0 = h->nparts;
  for (int s = 0; s < nsteps; ++s) {
    int rc = launch_step(h, dt, dt_mode, 1);
    if (rc != PPK_OK) return rc;
    h->iter_host += 1;
  }
  int rc = read_ctrl(h);
  if (rc != PPK_OK) return rc;
  h->iter_host = (long long)h->ctrl_host[ppk::CTRL_ITER];
  // no step ran (already finished): the CFL partials are as they were before the call
  if (h->iter_host == start) {
    h->cfl_valid = cfl_valid0;
    h->nparts = nparts0;
  }
  return PPK_OK;
}

int ppk_mhd_get_time(ppk_mhd_t *h, double *t, double *dt, int64_t *iteration) {
  if (!h) return fail(h, PPK_ERR_INVALID_ARGUMENT, "NULL handle");
  int rc = read_ctrl(h);
  if (rc != PPK_OK) return rc;
  if (t) *t = h->ctrl_host[ppk::CTRL_T];
  if (dt) *dt = h->ctrl_host[ppk::CTRL_DT];
  if (iteration) *iteration = (int64_t)h->ctrl_host[ppk::CTRL_ITER];
  return PPK_OK;
}

int ppk_mhd_set_time(ppk_mhd_t *h, double t, int64_t iteration) {
  if (!h) return fail(h, PPK_ERR_INVALID_ARGUMENT, "NULL handle");
  if (iteration < 0) return fail(h, PPK_ERR_INVALID_ARGUMENT, "iteration must be >= 0");
  int rc = read_ctrl(h);
  if (rc != PPK_OK) return rc;
  // keep the current data where it is: swap the ping-pong buffers if the parity changes
  if (((long long)iteration & 1) != (h->iter_host & 1)) std::swap(h->U[0], h->U[1]);
  h->cfl_valid = false;
  h->iter_host = iteration;
  h->ctrl_host[ppk::CTRL_T] = t;
  h->ctrl_host[ppk::CTRL_ITER] = (double)iteration;
  return write_ctrl(h);
}

int ppk_mhd_finished(ppk_mhd_t *h, int *finished) {
  if (!h || !finished) return fail(h, PPK_ERR_INVALID_ARGUMENT, "NULL argument");
  int rc = read_ctrl(h);
  if (rc != PPK_OK) return rc;
  // SolverBase::finished, src/shared/SolverBase.cpp:195-201
  *finished = (h->ctrl_host[ppk::CTRL_T] >= (h->p.tEnd - 1e-14) ||
               h->ctrl_host[ppk::CTRL_ITER] >= (double)h->p.n_step_max)
                  ? 1
                  : 0;
  return PPK_OK;
}

int ppk_mhd_synchronize(ppk_mhd_t *h) {
  if (!h) return fail(h, PPK_ERR_INVALID_ARGUMENT, "NULL handle");
  HIP_CHECK(h, hipStreamSynchronize(h->stream));
  return PPK_OK;
}

int ppk_mhd_timers_enable(ppk_mhd_t *h, uint32_t stage_mask) {
  if (!h) return fail(h, PPK_ERR_INVALID_ARGUMENT, "NULL handle");
  HIP_CHECK(h, hipStreamSynchronize(h->stream));
  h->timer->set_mask(stage_mask);
  return PPK_OK;
}

int ppk_mhd_timers_read(ppk_mhd_t *h, int stage, double *total_ms, int64_t *launches) {
  if (!h) return fail(h, PPK_ERR_INVALID_ARGUMENT, "NULL handle");
  if (stage < 0 || stage >= ppk::ST_COUNT) return fail(h, PPK_ERR_INVALID_ARGUMENT, "bad stage");
  HIP_CHECK(h, h->timer->resolve());
  if (total_ms) *total_ms = h->timer->total(stage);
  if (launches) *launches = h->timer->count(stage);
  return PPK_OK;
}

int ppk_mhd_timer_count(void) { return ppk::ST_COUNT; }

const char *ppk_mhd_timer_name(int stage) {
  if (stage < 0 || stage >= ppk::ST_COUNT) return "";
  return ppk::kStageNames[stage];
}

int ppk_mhd_divb_max(ppk_mhd_t *h, double *divb_max) {
  if (!h || !divb_max) return fail(h, PPK_ERR_INVALID_ARGUMENT, "NULL argument");
  int rc = ppk_mhd_make_boundaries(h);
  if (rc != PPK_OK) return rc;
  HIP_CHECK(h, ppk::diag_divb(h->dp, h->cur(), h->diag, h->stream));
  HIP_CHECK(h, hipMemcpyAsync(h->ctrl_host + ppk::CTRL_SLOTS, h->diag, sizeof(double),
                              hipMemcpyDeviceToHost, h->stream));
  HIP_CHECK(h, hipStreamSynchronize(h->stream));
  *divb_max = h->ctrl_host[ppk::CTRL_SLOTS];
  return PPK_OK;
}

int ppk_mhd_save_vtk(ppk_mhd_t *h, const char *dir, const char *prefix, int istep, int ascii) {
  if (!h || !dir || !prefix) return fail(h, PPK_ERR_INVALID_ARGUMENT, "NULL argument");
  std::vector<double> host((size_t)h->cells() * 8);
  int rc = ppk_mhd_download(h, host.data(), host.size());
  if (rc != PPK_OK) return rc;
  rc = ppk_mhd_write_vtk(&h->p, host.data(), dir, prefix, istep, ascii);
  if (rc != PPK_OK) return fail(h, rc, ppk::thread_error());
  return PPK_OK;
}

int ppk_mhd_conserved_sums(ppk_mhd_t *h, double sums[8]) {
  if (!h || !sums) return fail(h, PPK_ERR_INVALID_ARGUMENT, "NULL argument");
  HIP_CHECK(h, ppk::diag_sums(h->dp, h->cur(), h->diag, ppk::kDiagBlocks, h->stream));
  std::vector<double> part((size_t)ppk::kDiagBlocks * 8);
  HIP_CHECK(h, hipMemcpyAsync(part.data(), h->diag, part.size() * sizeof(double),
                              hipMemcpyDeviceToHost, h->stream));
  HIP_CHECK(h, hipStreamSynchronize(h->stream));
  for (int v = 0; v < 8; ++v) {
    double s = 0.0;
    for (int b = 0; b < ppk::kDiagBlocks; ++b) s += part[(size_t)b * 8 + v];
    sums[v] = s;
  }
  return PPK_OK;
}

int ppk_mhd_math_selfcheck(const double *x, int64_t n, double *out) {
  if (n < 0 || (n > 0 && (!x || !out)))
    return fail(nullptr, PPK_ERR_INVALID_ARGUMENT, "NULL array or negative count");
  if (n == 0) return PPK_OK;
  const size_t bytes = (size_t)n * sizeof(double);
  double *dx = nullptr, *dout = nullptr;
  hipError_t e = hipMalloc(&dx, bytes);
  if (e == hipSuccess) e = hipMalloc(&dout, 3 * bytes);
  if (e == hipSuccess) e = hipMemcpy(dx, x, bytes, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = ppk::diag_math(dx, n, dout, nullptr);
  if (e == hipSuccess) e = hipMemcpy(out, dout, 3 * bytes, hipMemcpyDeviceToHost);
  if (dx) (void)hipFree(dx);
  if (dout) (void)hipFree(dout);
  if (e != hipSuccess)
    return fail(nullptr, PPK_ERR_HIP, std::string("math self-check: ") + hipGetErrorString(e));
  return PPK_OK;
}

}  // extern "C"
