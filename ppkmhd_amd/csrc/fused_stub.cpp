// Placeholder until the fused kernels are built: the staged kernels serve every request.
#include "fused.hpp"

namespace ppk {
bool fused_supported(const DevParams &, int, int) { return false; }
hipError_t fused_alloc(const DevParams &, int, FusedScratch &) { return hipErrorNotSupported; }
void fused_free(FusedScratch &) {}
hipError_t fused_step(const DevParams &, double *, double *, FusedScratch &, double *, double, int,
                      int, hipStream_t, StageTimer *) {
  return hipErrorNotSupported;
}
hipError_t fused_cfl(const DevParams &, const double *, FusedScratch &, double *, hipStream_t,
                     StageTimer *) {
  return hipErrorNotSupported;
}
}  // namespace ppk
