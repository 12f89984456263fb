// Launchers of the surrogate network's kernels (mlp.hip) and of the surrogate face kernels
// (surrogate.hip), for the host units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "mlp.hpp"

namespace ppk {
namespace mlp {

// bf16 elements of the per-minibatch scratch of train_step
size_t fb_scratch_elems(long long B);
// One Adam step on a minibatch of B samples (B % 256 == 0) drawn from the n samples X, T
// (idx: explicit sample indices, or null to draw them from `seed` and `step`).  partials:
// (B / 256) x 19720 floats.  grad_out (optional): the summed gradient.  lr == 0 computes the
// gradient only and leaves the parameters unchanged.
hipError_t train_step(const MlpPacked *net, float *P, float *m, float *v, const float *X,
                      const float *T, const long long *idx, long long n, unsigned long long seed,
                      long long step, long long B, uint16_t *scratch, float *partials,
                      float *grad_out, float lr, hipStream_t s);
hipError_t pack(const float *P, MlpPacked *net, hipStream_t s);
int infer_blocks(long long n);
// standardised outputs Y[n][8] (optional) and per-block squared-error sums against T
// (optional, infer_blocks(n) x 8 doubles)
hipError_t infer(const MlpPacked *net, const float *X, const float *T, long long n, float *Y,
                 double *sq_parts, hipStream_t s);
// per-block sums and sums of squares of the 24 columns: nblocks x 24 x 2 doubles
hipError_t moments(const float *X, const float *T, long long n, double *parts, int nblocks,
                   hipStream_t s);

}  // namespace mlp
}  // namespace ppk
