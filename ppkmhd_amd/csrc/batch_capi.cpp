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
  long long running[ppk::BS_COUNT] = {};
  for (int m = 0; m < nmembers; ++m) {
    ppk::BatchMember &bm = h->members[m];
    std::memset(&bm, 0, sizeof(bm));
    bm.p = ppk::make_dev_params(params[m]);
    bm.p.bulk32 = h->bulk32;
    bm.p.brick = 0;
    bm.p.tile3 = 0;
    bm.p.riemann_mix = 1;
    const size_t n = (size_t)bm.p.vstride;
    const int d = bm.p.dim;
    bm.U[0] = static_cast<double *>(dalloc(8 * n * sizeof(double)));
    bm.U[1] = static_cast<double *>(dalloc(8 * n * sizeof(double)));
    bm.rec = dalloc((size_t)(d == 3 ? ppk::kRec3 : ppk::kRec2) * n * es);
    bm.flux = dalloc((size_t)d * (d == 3 ? ppk::kFluxComp3 : ppk::kFluxComp2) * n * es);
    bm.emf = dalloc((size_t)(d == 3 ? 3 : 1) * n * es);
    bm.pad = ppk::padded_bricks(bm.p);
    bm.sub = ppk::sub_bricks(bm.p);
    bm.nparts = ppk::fused_cfl_parts(bm.p);
    bm.parts = static_cast<double *>(dalloc(sizeof(double) * bm.nparts));
    bm.ctrl = h->ctrl + (size_t)m * ppk::CTRL_SLOTS;
    if (!bm.U[0] || !bm.U[1] || !bm.rec || !bm.flux || !bm.emf || !bm.parts)
      return bail(PPK_ERR_OUT_OF_MEMORY, "device allocation failed (member " +
                                             std::to_string(m) + ")");
    if (hipMemsetAsync(bm.U[0], 0, 8 * n * sizeof(double), h->stream) != hipSuccess ||
        hipMemsetAsync(bm.U[1], 0, 8 * n * sizeof(double), h->stream) != hipSuccess)
      return bail(PPK_ERR_HIP, "hipMemsetAsync failed");
    long long counts[ppk::BS_COUNT];
    h->K->batch_block_counts(bm.p, counts);
    for (int st = 0; st < ppk::BS_COUNT; ++st) {
      bm.begin[st] = running[st];
      running[st] += counts[st];
    }
    double *c = h->ctrl_host + (size_t)m * ppk::CTRL_SLOTS;
    c[ppk::CTRL_DT] = params[m].tEnd;
    c[ppk::CTRL_TEND] = params[m].tEnd;
    c[ppk::CTRL_NSTEPMAX] = (double)params[m].n_step_max;
    c[ppk::CTRL_ACTIVE] = 1.0;
  }
  for (int st = 0; st < ppk::BS_COUNT; ++st) {
    if (running[st] > 0x7fffffffLL) return bail(PPK_ERR_INVALID_ARGUMENT, "batch too large");
    h->totals[st] = running[st];
  }
  h->d_members =
      static_cast<ppk::BatchMember *>(dalloc(sizeof(ppk::BatchMember) * (size_t)nmembers));
  if (!h->d_members) return bail(PPK_ERR_OUT_OF_MEMORY, "device allocation failed");
  if (hipMemcpyAsync(h->d_members, h->members.data(), sizeof(ppk::BatchMember) * nmembers,
                     hipMemcpyHostToDevice, h->stream) != hipSuccess)
    return bail(PPK_ERR_HIP, "member table upload failed");
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
  int rc = read_ctrl(h);
  if (rc != PPK_OK) return rc;
  for (int m = 0; m < h->nm; ++m) {
    const ppk::BatchMember &bm = h->members[m];
    std::vector<double> host((size_t)bm.p.vstride * 8);
    ppk::init_problem_host(h->params[m], problems[m], host.data());
    BHIP(h, hipMemcpyAsync(bm.U[0], host.data(), host.size() * sizeof(double),
                           hipMemcpyHostToDevice, h->stream));
    BHIP(h, hipStreamSynchronize(h->stream));
    double *c = h->ctrl_host + (size_t)m * ppk::CTRL_SLOTS;
    c[ppk::CTRL_T] = 0.0;
    c[ppk::CTRL_ITER] = 0.0;
    c[ppk::CTRL_DT] = h->params[m].tEnd;
    c[ppk::CTRL_ACTIVE] = 1.0;
  }
  h->cfl_valid = false;
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
    BHIP(h, h->K->batch_step(h->d_members, h->nm, h->dim, h->bulk32, h->totals, 1, h->stream));
    h->cfl_valid = true;
    done = 1;
  }
  if (done < nsteps && !h->exec) {
    BHIP(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    hipError_t e = h->K->batch_step(h->d_members, h->nm, h->dim, h->bulk32, h->totals, 0,
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

int ppk_mhd_batch_synchronize(ppk_mhd_batch_t *h) {
  if (!h) return bfail(h, PPK_ERR_INVALID_ARGUMENT, "NULL handle");
  BHIP(h, hipStreamSynchronize(h->stream));
  return PPK_OK;
}

}  // extern "C"
