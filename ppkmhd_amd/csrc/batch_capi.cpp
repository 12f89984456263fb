// C ABI of batched ensembles (include/ppkmhd_amd.h, ppk_mhd_batch_*): many independent grids
// advanced together, one launch per stage for all of them, one hipGraph per step.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "capi_internal.hpp"
#include "kernels.hpp"
#include "ppkmhd_amd.h"

struct ppk_mhd_batch {
  int nm = 0, dim = 0, bulk32 = 0;
  int tile = 0, brick = 0;  // half-step tile shape and brick shape of the block tables
  const ppk::KernelTable *K = nullptr;
  std::vector<ppk_mhd_params> params;
  std::vector<ppk::BatchMember> members;  // host copy of the device table
  ppk::BatchMember *d_members = nullptr;
  double *ctrl = nullptr;       // nm control blocks (device)
  double *ctrl_host = nullptr;  // pinned mirror
  long long totals[ppk::BS_COUNT] = {};
  std::vector<void *> allocs;
  hipStream_t stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool cfl_valid = false;
  std::string err;
};

namespace {

int bfail(ppk_mhd_batch_t *h, int code, const std::string &msg) {
  if (h)
    h->err = msg;
  else
    ppk::set_thread_error(msg);
  return code;
}

#define BHIP(h, expr)                                                                     \
  do {                                                                                    \
    hipError_t _e = (expr);                                                               \
    if (_e != hipSuccess)                                                                 \
      return bfail((h), PPK_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e));  \
  } while (0)

void bfree(ppk_mhd_batch_t *h) {
  if (h->exec) (void)hipGraphExecDestroy(h->exec);
  if (h->graph) (void)hipGraphDestroy(h->graph);
  for (void *p : h->allocs) (void)hipFree(p);
  h->allocs.clear();
  if (h->ctrl_host) (void)hipHostFree(h->ctrl_host);
  if (h->stream) (void)hipStreamDestroy(h->stream);
}

int read_ctrl(ppk_mhd_batch_t *h) {
  BHIP(h, hipMemcpyAsync(h->ctrl_host, h->ctrl, sizeof(double) * ppk::CTRL_SLOTS * h->nm,
                         hipMemcpyDeviceToHost, h->stream));
  BHIP(h, hipStreamSynchronize(h->stream));
  return PPK_OK;
}
int write_ctrl(ppk_mhd_batch_t *h) {
  BHIP(h, hipMemcpyAsync(h->ctrl, h->ctrl_host, sizeof(double) * ppk::CTRL_SLOTS * h->nm,
                         hipMemcpyHostToDevice, h->stream));
  BHIP(h, hipStreamSynchronize(h->stream));
  return PPK_OK;
}

// brick shapes of the update / face / edge block tables (make_brick): 2D has two distinct ones
constexpr int kBrickShapes = 4;
int brick_shapes(int dim) { return dim == 3 ? kBrickShapes : 2; }

void drop_graph(ppk_mhd_batch_t *h) {
  if (h->exec) (void)hipGraphExecDestroy(h->exec);
  if (h->graph) (void)hipGraphDestroy(h->graph);
  h->exec = nullptr;
  h->graph = nullptr;
}

// block tables of every member for a tile / brick shape: first blocks per stage, totals, the
// bricks; uploads the member table and drops the captured step (its grid sizes change)
int set_tables(ppk_mhd_batch_t *h, int tile, int brick) {
  if (tile < 0 || tile >= ppk::kBatchTiles || brick < 0 || brick >= brick_shapes(h->dim))
    return bfail(h, PPK_ERR_INVALID_ARGUMENT, "tile or brick shape out of range");
  long long running[ppk::BS_COUNT] = {};
  for (ppk::BatchMember &bm : h->members) {
    bm.p.brick = brick;
    bm.pad = ppk::padded_bricks(bm.p);
    bm.sub = ppk::sub_bricks(bm.p);
    bm.nparts = ppk::fused_cfl_parts(bm.p);
    long long counts[ppk::BS_COUNT];
    h->K->batch_block_counts(bm.p, tile, counts);
    for (int st = 0; st < ppk::BS_COUNT; ++st) {
      bm.begin[st] = running[st];
      running[st] += counts[st];
    }
  }
  for (int st = 0; st < ppk::BS_COUNT; ++st) {
    if (running[st] > 0x7fffffffLL) return bfail(h, PPK_ERR_INVALID_ARGUMENT, "batch too large");
    h->totals[st] = running[st];
  }
  BHIP(h, hipStreamSynchronize(h->stream));
  BHIP(h, hipMemcpyAsync(h->d_members, h->members.data(), sizeof(ppk::BatchMember) * h->nm,
                         hipMemcpyHostToDevice, h->stream));
  BHIP(h, hipStreamSynchronize(h->stream));
  drop_graph(h);
  h->tile = tile;
  h->brick = brick;
  h->cfl_valid = false;  // the partials' layout follows the brick shape
  return PPK_OK;
}

// the step graph (one batch_step without the standalone CFL pass), captured once per table
int capture_step(ppk_mhd_batch_t *h) {
  if (h->exec) return PPK_OK;
  BHIP(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
  hipError_t e = h->K->batch_step(h->d_members, h->nm, h->dim, h->bulk32, h->tile, h->totals, 0,
                                  h->stream);
  hipGraph_t g = nullptr;
  hipError_t e2 = hipStreamEndCapture(h->stream, &g);
  if (e != hipSuccess || e2 != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    return bfail(h, PPK_ERR_HIP, std::string("step graph capture failed: ") +
                                     hipGetErrorString(e != hipSuccess ? e : e2));
  }
  h->graph = g;
  BHIP(h, hipGraphInstantiate(&h->exec, h->graph, nullptr, nullptr, 0));
  return PPK_OK;
}

int check_member(ppk_mhd_batch_t *h, int m) {
  if (m < 0 || m >= h->nm) return bfail(h, PPK_ERR_INVALID_ARGUMENT, "member index out of range");
  return PPK_OK;
}

}  // namespace

extern "C" {

int ppk_mhd_batch_create(const ppk_mhd_params *params, int nmembers, ppk_mhd_batch_t **handle) {
  if (!handle || !params) return bfail(nullptr, PPK_ERR_INVALID_ARGUMENT, "NULL argument");
  *handle = nullptr;
  if (nmembers < 1) return bfail(nullptr, PPK_ERR_INVALID_ARGUMENT, "nmembers must be >= 1");
  for (int m = 0; m < nmembers; ++m) {
    std::string why;
    int rc = ppk::validate_params(&params[m], why);
    if (rc != PPK_OK) return bfail(nullptr, rc, "member " + std::to_string(m) + ": " + why);
    if (params[m].dim != params[0].dim || params[m].precision != params[0].precision)
      return bfail(nullptr, PPK_ERR_INVALID_ARGUMENT,
                   "all members of a batch share dim and precision");
  }
  ppk_mhd_batch_t *h = new (std::nothrow) ppk_mhd_batch_t();
  if (!h) return bfail(nullptr, PPK_ERR_OUT_OF_MEMORY, "host allocation failed");
  auto bail = [&](int code, const std::string &msg) {
    bfree(h);
    delete h;
    return bfail(nullptr, code, msg);
  };
  h->nm = nmembers;
  h->dim = params[0].dim;
  h->bulk32 = params[0].precision == PPK_FP32_BULK ? 1 : 0;
  h->K = params[0].precision == PPK_FP64_STRICT ? &ppk::kStrict : &ppk::kFast;
  h->params.assign(params, params + nmembers);
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess)
    return bail(PPK_ERR_HIP, "hipStreamCreateWithFlags failed");
  auto dalloc = [&](size_t bytes) -> void * {
    void *p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
    h->allocs.push_back(p);
    return p;
  };
  h->ctrl = static_cast<double *>(dalloc(sizeof(double) * ppk::CTRL_SLOTS * nmembers));
  if (!h->ctrl) return bail(PPK_ERR_OUT_OF_MEMORY, "device allocation failed");
  if (hipHostMalloc((void **)&h->ctrl_host, sizeof(double) * ppk::CTRL_SLOTS * nmembers) !=
      hipSuccess)
    return bail(PPK_ERR_OUT_OF_MEMORY, "pinned host allocation failed");
  std::fill(h->ctrl_host, h->ctrl_host + ppk::CTRL_SLOTS * nmembers, 0.0);
  const size_t es = h->bulk32 ? sizeof(float) : sizeof(double);
  h->members.resize(nmembers);
  for (int m = 0; m < nmembers; ++m) {
    ppk::BatchMember &bm = h->members[m];
    std::memset(&bm, 0, sizeof(bm));
    bm.p = ppk::make_dev_params(params[m]);
    bm.p.bulk32 = h->bulk32;
    bm.p.brick = 0;
    bm.p.tile3 = 0;
    const size_t n = (size_t)bm.p.vstride;
    const int d = bm.p.dim;
    bm.U[0] = static_cast<double *>(dalloc(8 * n * sizeof(double)));
    bm.U[1] = static_cast<double *>(dalloc(8 * n * sizeof(double)));
    bm.rec = dalloc((size_t)(d == 3 ? ppk::kRec3 : ppk::kRec2) * n * es);
    bm.flux = dalloc((size_t)d * (d == 3 ? ppk::kFluxComp3 : ppk::kFluxComp2) * n * es);
    bm.emf = dalloc((size_t)(d == 3 ? 3 : 1) * n * sizeof(double));  // FP64 EMFs in every form
    int most_parts = 0;  // the partials of the largest brick table any shape gives
    for (int shape = 0; shape < kBrickShapes; ++shape) {
      ppk::DevParams q = bm.p;
      q.brick = shape;
      most_parts = std::max(most_parts, ppk::fused_cfl_parts(q));
    }
    bm.parts = static_cast<double *>(dalloc(sizeof(double) * most_parts));
    bm.ctrl = h->ctrl + (size_t)m * ppk::CTRL_SLOTS;
    if (!bm.U[0] || !bm.U[1] || !bm.rec || !bm.flux || !bm.emf || !bm.parts)
      return bail(PPK_ERR_OUT_OF_MEMORY, "device allocation failed (member " +
                                             std::to_string(m) + ")");
    if (hipMemsetAsync(bm.U[0], 0, 8 * n * sizeof(double), h->stream) != hipSuccess ||
        hipMemsetAsync(bm.U[1], 0, 8 * n * sizeof(double), h->stream) != hipSuccess)
      return bail(PPK_ERR_HIP, "hipMemsetAsync failed");
    double *c = h->ctrl_host + (size_t)m * ppk::CTRL_SLOTS;
    c[ppk::CTRL_DT] = params[m].tEnd;
    c[ppk::CTRL_TEND] = params[m].tEnd;
    c[ppk::CTRL_NSTEPMAX] = (double)params[m].n_step_max;
    c[ppk::CTRL_ACTIVE] = 1.0;
  }
  h->d_members =
      static_cast<ppk::BatchMember *>(dalloc(sizeof(ppk::BatchMember) * (size_t)nmembers));
  if (!h->d_members) return bail(PPK_ERR_OUT_OF_MEMORY, "device allocation failed");
  if (set_tables(h, 0, 0) != PPK_OK) return bail(PPK_ERR_INVALID_ARGUMENT, h->err);
  if (write_ctrl(h) != PPK_OK) return bail(PPK_ERR_HIP, h->err);
  *handle = h;
  return PPK_OK;
}

void ppk_mhd_batch_destroy(ppk_mhd_batch_t *h) {
  if (!h) return;
  (void)hipStreamSynchronize(h->stream);
  bfree(h);
  delete h;
}

const char *ppk_mhd_batch_last_error(const ppk_mhd_batch_t *h) {
  return h ? h->err.c_str() : ppk::thread_error();
}

int ppk_mhd_batch_init_problem(ppk_mhd_batch_t *h, const ppk_mhd_problem *problems) {
  if (!h || !problems) return bfail(h, PPK_ERR_INVALID_ARGUMENT, "NULL argument");
  for (int m = 0; m < h->nm; ++m)
    if (!ppk::mhd_problem_ok(problems[m].kind))
      return bfail(h, PPK_ERR_INVALID_ARGUMENT, "unknown problem kind (member " +
                                                    std::to_string(m) + ")");
  int rc = read_ctrl(h);
  if (rc != PPK_OK) return rc;
  for (int m = 0; m < h->nm; ++m) {
    double *c = h->ctrl_host + (size_t)m * ppk::CTRL_SLOTS;
    c[ppk::CTRL_T] = 0.0;
    c[ppk::CTRL_ITER] = 0.0;
    c[ppk::CTRL_ACTIVE] = 1.0;
  }
  rc = write_ctrl(h);
  if (rc != PPK_OK) return rc;
  // as a single solver's init: the state, its ghost fill, U2 = U, and the CFL dt of t = 0
  for (int m = 0; m < h->nm; ++m) {
    const ppk::BatchMember &bm = h->members[m];
    std::vector<double> host((size_t)bm.p.vstride * 8);
    ppk::init_problem_host(h->params[m], problems[m], host.data());
    BHIP(h, hipMemcpyAsync(bm.U[0], host.data(), host.size() * sizeof(double),
                           hipMemcpyHostToDevice, h->stream));
    BHIP(h, h->K->boundaries(bm.p, bm.U[0], nullptr, bm.ctrl, h->stream, nullptr));
    BHIP(h, h->K->fused_cfl(bm.p, bm.U[0], bm.parts, bm.ctrl, h->stream, nullptr));
    BHIP(h, h->K->set_dt(bm.ctrl, bm.parts, bm.nparts, 0.0, ppk::SETDT_REDUCE_ONLY,
                         h->params[m].cfl, h->stream));
    BHIP(h, hipMemcpyAsync(bm.U[1], bm.U[0], host.size() * sizeof(double),
                           hipMemcpyDeviceToDevice, h->stream));
    BHIP(h, hipStreamSynchronize(h->stream));
  }
  rc = read_ctrl(h);
  if (rc != PPK_OK) return rc;
  for (int m = 0; m < h->nm; ++m) {
    double *c = h->ctrl_host + (size_t)m * ppk::CTRL_SLOTS;
    const double dt = h->params[m].cfl / c[ppk::CTRL_INVDT];
    c[ppk::CTRL_DT] = (0.0 + dt > h->params[m].tEnd) ? h->params[m].tEnd - 0.0 : dt;
  }
  h->cfl_valid = true;
  return write_ctrl(h);
}

int ppk_mhd_batch_upload(ppk_mhd_batch_t *h, int member, const double *U, size_t count) {
  if (!h || !U) return bfail(h, PPK_ERR_INVALID_ARGUMENT, "NULL argument");
  int rc = check_member(h, member);
  if (rc != PPK_OK) return rc;
  const ppk::BatchMember &bm = h->members[member];
  if (count != (size_t)bm.p.vstride * 8)
    return bfail(h, PPK_ERR_INVALID_ARGUMENT, "count must be 8*isize*jsize*ksize");
  rc = read_ctrl(h);
  if (rc != PPK_OK) return rc;
  const int par = ((long long)h->ctrl_host[(size_t)member * ppk::CTRL_SLOTS + ppk::CTRL_ITER]) & 1;
  BHIP(h, hipMemcpyAsync(bm.U[par], U, count * sizeof(double), hipMemcpyHostToDevice, h->stream));
  BHIP(h, hipStreamSynchronize(h->stream));
  h->cfl_valid = false;
  return PPK_OK;
}

int ppk_mhd_batch_download(ppk_mhd_batch_t *h, int member, double *U, size_t count) {
  if (!h || !U) return bfail(h, PPK_ERR_INVALID_ARGUMENT, "NULL argument");
  int rc = check_member(h, member);
  if (rc != PPK_OK) return rc;
  const ppk::BatchMember &bm = h->members[member];
  if (count != (size_t)bm.p.vstride * 8)
    return bfail(h, PPK_ERR_INVALID_ARGUMENT, "count must be 8*isize*jsize*ksize");
  rc = read_ctrl(h);
  if (rc != PPK_OK) return rc;
  const int par = ((long long)h->ctrl_host[(size_t)member * ppk::CTRL_SLOTS + ppk::CTRL_ITER]) & 1;
  BHIP(h, hipMemcpyAsync(U, bm.U[par], count * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  BHIP(h, hipStreamSynchronize(h->stream));
  return PPK_OK;
}

int ppk_mhd_batch_run(ppk_mhd_batch_t *h, int nsteps) {
  if (!h) return bfail(h, PPK_ERR_INVALID_ARGUMENT, "NULL handle");
  if (nsteps < 0) return bfail(h, PPK_ERR_INVALID_ARGUMENT, "nsteps must be >= 0");
  int done = 0;
  if (nsteps > 0 && !h->cfl_valid) {
    // first step: CFL of the current states from a standalone pass
    BHIP(h, h->K->batch_step(h->d_members, h->nm, h->dim, h->bulk32, h->tile, h->totals, 1,
                                  h->stream));
    h->cfl_valid = true;
    done = 1;
  }
  if (done < nsteps) {
    const int rc = capture_step(h);
    if (rc != PPK_OK) return rc;
  }
  for (; done < nsteps; ++done) BHIP(h, hipGraphLaunch(h->exec, h->stream));
  BHIP(h, hipStreamSynchronize(h->stream));
  return PPK_OK;
}

int ppk_mhd_batch_get_time(ppk_mhd_batch_t *h, int member, double *t, double *dt,
                           int64_t *iteration) {
  if (!h) return bfail(h, PPK_ERR_INVALID_ARGUMENT, "NULL handle");
  int rc = check_member(h, member);
  if (rc != PPK_OK) return rc;
  rc = read_ctrl(h);
  if (rc != PPK_OK) return rc;
  const double *c = h->ctrl_host + (size_t)member * ppk::CTRL_SLOTS;
  if (t) *t = c[ppk::CTRL_T];
  if (dt) *dt = c[ppk::CTRL_DT];
  if (iteration) *iteration = (int64_t)c[ppk::CTRL_ITER];
  return PPK_OK;
}

int ppk_mhd_batch_size(const ppk_mhd_batch_t *h) { return h ? h->nm : 0; }

int ppk_mhd_batch_set_tiles(ppk_mhd_batch_t *h, int tile, int brick) {
  if (!h) return bfail(h, PPK_ERR_INVALID_ARGUMENT, "NULL handle");
  return set_tables(h, tile, brick);
}

int ppk_mhd_batch_get_tiles(const ppk_mhd_batch_t *h, int *tile, int *brick) {
  if (!h) return bfail(nullptr, PPK_ERR_INVALID_ARGUMENT, "NULL handle");
  if (tile) *tile = h->tile;
  if (brick) *brick = h->brick;
  return PPK_OK;
}

int ppk_mhd_batch_autotune(ppk_mhd_batch_t *h, int trial_steps, int *tile, int *brick,
                           double *ms_per_step) {
  if (!h) return bfail(h, PPK_ERR_INVALID_ARGUMENT, "NULL handle");
  if (trial_steps < 1) return bfail(h, PPK_ERR_INVALID_ARGUMENT, "trial_steps must be >= 1");
  // snapshot: control blocks and every member's current state
  int rc = read_ctrl(h);
  if (rc != PPK_OK) return rc;
  std::vector<double> ctrl_saved(h->ctrl_host, h->ctrl_host + (size_t)ppk::CTRL_SLOTS * h->nm);
  size_t total = 0;
  for (const ppk::BatchMember &bm : h->members) total += (size_t)bm.p.vstride * 8;
  double *backup = nullptr;
  BHIP(h, hipMalloc(&backup, total * sizeof(double)));
  auto current = [&](int m) {
    const long long it = (long long)ctrl_saved[(size_t)m * ppk::CTRL_SLOTS + ppk::CTRL_ITER];
    return h->members[m].U[it & 1];
  };
  auto copy_states = [&](bool save) -> hipError_t {
    size_t off = 0;
    for (int m = 0; m < h->nm; ++m) {
      const size_t n = (size_t)h->members[m].p.vstride * 8;
      hipError_t e = save ? hipMemcpyAsync(backup + off, current(m), n * sizeof(double),
                                           hipMemcpyDeviceToDevice, h->stream)
                          : hipMemcpyAsync(current(m), backup + off, n * sizeof(double),
                                           hipMemcpyDeviceToDevice, h->stream);
      if (e != hipSuccess) return e;
      off += n;
    }
    return hipStreamSynchronize(h->stream);
  };
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int best_tile = h->tile, best_brick = h->brick;
  double best_ms = -1.0;
  int status = PPK_OK;
  do {
    if (copy_states(true) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
        hipEventCreate(&e1) != hipSuccess) {
      status = bfail(h, PPK_ERR_HIP, "autotune: snapshot failed");
      break;
    }
    for (int tl = 0; tl < ppk::kBatchTiles && status == PPK_OK; ++tl)
      for (int bk = 0; bk < brick_shapes(h->dim) && status == PPK_OK; ++bk) {
        if ((status = set_tables(h, tl, bk)) != PPK_OK) break;
        // one step outside the timing (standalone CFL), the graph capture, then the replays
        if ((status = ppk_mhd_batch_run(h, 1)) != PPK_OK) break;
        if ((status = capture_step(h)) != PPK_OK) break;
        float ms = 0.f;
        if (hipEventRecord(e0, h->stream) != hipSuccess) status = PPK_ERR_HIP;
        for (int s = 0; s < trial_steps && status == PPK_OK; ++s)
          if (hipGraphLaunch(h->exec, h->stream) != hipSuccess) status = PPK_ERR_HIP;
        if (status == PPK_OK &&
            (hipEventRecord(e1, h->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
             hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
          status = PPK_ERR_HIP;
        if (status != PPK_OK) {
          bfail(h, status, "autotune: trial launch failed");
          break;
        }
        const double per = ms / trial_steps;
        if (best_ms < 0.0 || per < best_ms) {
          best_ms = per;
          best_tile = tl;
          best_brick = bk;
        }
        // back to the snapshot
        std::copy(ctrl_saved.begin(), ctrl_saved.end(), h->ctrl_host);
        if (write_ctrl(h) != PPK_OK || copy_states(false) != hipSuccess) {
          status = bfail(h, PPK_ERR_HIP, "autotune: restore failed");
          break;
        }
      }
  } while (false);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipFree(backup);
  if (status != PPK_OK) return status;
  rc = set_tables(h, best_tile, best_brick);
  if (rc != PPK_OK) return rc;
  if (tile) *tile = best_tile;
  if (brick) *brick = best_brick;
  if (ms_per_step) *ms_per_step = best_ms;
  return PPK_OK;
}

int ppk_mhd_batch_synchronize(ppk_mhd_batch_t *h) {
  if (!h) return bfail(h, PPK_ERR_INVALID_ARGUMENT, "NULL handle");
  BHIP(h, hipStreamSynchronize(h->stream));
  return PPK_OK;
}

}  // extern "C"
