// Internal declarations shared by the host-side translation units of the library.
#pragma once

#include <string>

#include "ppkmhd_amd.h"

namespace ppk {

// message of the last failure of a call that has no handle (create, .ini parsing), per thread
void set_thread_error(const std::string &msg);
const char *thread_error();

// parameter checks and device-side parameter block of a solver (capi.cpp)
int validate_params(const ppk_mhd_params *p, std::string &why);
struct DevParams;
DevParams make_dev_params(const ppk_mhd_params &p);

// host initial conditions in the reference layout (ics.cpp)
void init_problem_host(const ppk_mhd_params &p, const ppk_mhd_problem &pr, double *U);

}  // namespace ppk
