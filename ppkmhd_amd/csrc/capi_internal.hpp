// Internal declarations shared by the host-side translation units of the library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "mlp.hpp"
#include "ppkmhd_amd.h"

namespace ppk {

// message of the last failure of a call that has no handle (create, .ini parsing), per thread
void set_thread_error(const std::string &msg);
const char *thread_error();

// parameter checks and device-side parameter block of a solver (capi.cpp)
int validate_params(const ppk_mhd_params *p, std::string &why);
bool mhd_problem_ok(int kind);
struct DevParams;
DevParams make_dev_params(const ppk_mhd_params &p);
// a solver handle's parameters, and its per-handle error message (capi.cpp)
ppk_mhd_params handle_params(const ppk_mhd_t *h);
void set_handle_error(ppk_mhd_t *h, const std::string &msg);

// host initial conditions in the reference layout (ics.cpp)
void init_problem_host(const ppk_mhd_params &p, const ppk_mhd_problem &pr, double *U);
// the same on the device, into U (8 x cells); A: 3 x cells of scratch (ics_device.hip)
hipError_t init_problem_device(const ppk_mhd_params &p, const ppk_mhd_problem &pr, double *U,
                               double *A, hipStream_t s);

}  // namespace ppk

// the surrogate network: FP32 master parameters, Adam moments, bf16 packed copy (device)
struct ppk_mlp {
  float *P = nullptr, *m = nullptr, *v = nullptr;
  ppk::MlpPacked *packed = nullptr;
  long long adam_step = 0;
  // training scratch, sized for the last minibatch
  uint16_t *scratch = nullptr;
  float *partials = nullptr;
  long long scratch_batch = 0;
  hipStream_t stream = nullptr;
  std::string err;
};

// training samples of the surrogate: features X[n][16], targets T[n][8] (device)
struct ppk_faceset {
  float *X = nullptr, *T = nullptr;
  size_t n = 0, capacity = 0;
  std::string err;
};
