// The CFL time step of a caller-owned device state with no device allocation
// (ppk_mhd_compute_dt_noalloc): the same ghost-fill, CFL and reduction kernels as
// ppk_mhd_compute_dt, with their control block and per-block partial maxima in a file-scope
// device array instead of a solver handle's buffers.
//
// The array exists once per device (each GPU the library runs on holds it until the process
// ends).  Its device address is looked up once per device, with that device current, and only
// that address is handed to kernels: in host code the array's name is not a device address.
// Every call on a device shares the array and leaves its contents behind; the kernels write
// every slot they later read (the control block's active flag by the begin kernel, each
// partial by its CFL block, the maximum by the reduction), so nothing needs clearing between
// calls.  Two calls that overlap on one device (two streams, two threads) race on it.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "capi_internal.hpp"
#include "kernels.hpp"
#include "ppkmhd_amd.h"

// CFL partial maxima: one per brick of 256 padded cells, so grids up to 2^28 padded cells
constexpr int kNoAllocParts = 1 << 20;
constexpr long long kNoAllocDoubles = ppk::CTRL_SLOTS + (long long)kNoAllocParts;  // 8 MiB

// [0, CTRL_SLOTS): control block; then the partial maxima
static __device__ __align__(256) double scratch[kNoAllocDoubles];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mutex;
double *g_scratch[kMaxDevices] = {};

// device address of this device's array (looked up on the first call on the device)
int device_scratch(double **out, std::string &why) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) {
    why = std::string("hipGetDevice: ") + hipGetErrorString(e);
    return PPK_ERR_HIP;
  }
  if (dev < 0 || dev >= kMaxDevices) {
    why = "device index beyond the library's table of scratch addresses";
    return PPK_ERR_UNSUPPORTED;
  }
  std::lock_guard<std::mutex> lock(g_mutex);
  if (!g_scratch[dev]) {
    void *ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
      why = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
      return PPK_ERR_HIP;
    }
    g_scratch[dev] = static_cast<double *>(ptr);
  }
  *out = g_scratch[dev];
  return PPK_OK;
}

// dt = cfl / max_cells sum_d (c_f + |v_d|)/dx_d, as ppk_mhd_compute_dt returns it (the IEEE
// quotient; no tEnd clamp)
__global__ void k_dt_out(const double *ctrl, double cfl, double *dt) {
  if (threadIdx.x == 0) dt[0] = cfl / ctrl[ppk::CTRL_INVDT];
}

int fail(int code, const std::string &msg) {
  ppk::set_thread_error(msg);
  return code;
}

}  // namespace

extern "C" int ppk_mhd_compute_dt_noalloc(const ppk_mhd_params *params, double *dU, double *dt,
                                          void *stream) {
  std::string why;
  int rc = ppk::validate_params(params, why);
  if (rc != PPK_OK) return fail(rc, why);
  if (!dU || !dt) return fail(PPK_ERR_INVALID_ARGUMENT, "NULL device pointer");
  const ppk::DevParams dp = ppk::make_dev_params(*params);
  const long long nparts = ppk::padded_bricks(dp).count();
  if (nparts > kNoAllocParts)
    return fail(PPK_ERR_INVALID_ARGUMENT,
                "grid too large for ppk_mhd_compute_dt_noalloc: it has " +
                    std::to_string(nparts) + " bricks of 256 padded cells, the limit is " +
                    std::to_string(kNoAllocParts) + " (2^28 padded cells)");
  double *s = nullptr;
  rc = device_scratch(&s, why);
  if (rc != PPK_OK) return fail(rc, why);
  double *ctrl = s, *parts = s + ppk::CTRL_SLOTS;
  const ppk::KernelTable &K = params->precision == PPK_FP64_STRICT ? ppk::kStrict : ppk::kFast;
  hipStream_t st = static_cast<hipStream_t>(stream);
  hipError_t e = K.begin(ctrl, 0, st);
  if (e == hipSuccess) e = K.boundaries(dp, dU, nullptr, ctrl, st, nullptr);
  if (e == hipSuccess) e = K.fused_cfl(dp, dU, parts, ctrl, st, nullptr);
  if (e == hipSuccess)
    e = K.set_dt(ctrl, parts, (int)nparts, 0.0, ppk::SETDT_REDUCE_ONLY, params->cfl, st);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_dt_out, dim3(1), dim3(64), 0, st, ctrl, params->cfl, dt);
    e = hipGetLastError();
  }
  if (e != hipSuccess) return fail(PPK_ERR_HIP, std::string("launch: ") + hipGetErrorString(e));
  return PPK_OK;
}
