// The surrogate network (16 -> 128 -> 128 -> 8, ReLU): parameter layout, the bf16 operand
// packing of its weights for the MFMA kernels, and the host-side handles of the C ABI.
//
// Master parameters are FP32, in this order (the layout of ppk_mlp_get_params):
//   W1[128][16], b1[128], W2[128][128], b2[128], W3[8][128], b3[8]      (trained)
//   mu_in[16], isig_in[16], mu_out[8], sig_out[8]                       (standardisation)
// A forward pass on features x: xs = (x - mu_in) * isig_in; h1 = relu(W1 xs + b1);
// h2 = relu(W2 h1 + b2); y = W3 h2 + b3; the flux correction is mu_out + sig_out * y.
// On the GPU, xs, h1, h2 and the weights are rounded to bf16 (round to nearest even) and the
// products accumulate in FP32 on the matrix cores.
//
// MFMA operand packing.  The kernels keep activations transposed, features x samples, so the
// output of one layer (the accumulator layout of a 16x16 MFMA tile: lane l holds rows
// 4(l/16)+j, j = 0..3, of column l%16) is directly the B operand of the next (lane l holds
// the k-group l/16 of column l%16).  A 16x16x32 step consumes two 16-row output tiles, so its
// logical k index (group g = l/16, element e) is the feature
//     feat(q, g, e) = 32 q + (e < 4 ? 4 g + e : 16 + 4 g + e - 4)
// and the weight (A) operands are stored in that order.  The order inside one MFMA's k range
// does not change the sum, so any hardware assignment of (g, e) to k that is the same for A
// and B gives the same product.
#pragma once

#include <cstdint>

namespace ppk {

constexpr int kMlpIn = 16, kMlpHid = 128, kMlpOut = 8;

constexpr int kMlpW1 = 0;
constexpr int kMlpB1 = kMlpW1 + 128 * 16;
constexpr int kMlpW2 = kMlpB1 + 128;
constexpr int kMlpB2 = kMlpW2 + 128 * 128;
constexpr int kMlpW3 = kMlpB2 + 128;
constexpr int kMlpB3 = kMlpW3 + 8 * 128;
constexpr int kMlpTrained = kMlpB3 + 8;  // 19720 trained parameters
constexpr int kMlpMuIn = kMlpTrained;
constexpr int kMlpIsigIn = kMlpMuIn + 16;
constexpr int kMlpMuOut = kMlpIsigIn + 16;
constexpr int kMlpSigOut = kMlpMuOut + 8;
constexpr int kMlpParams = kMlpSigOut + 8;  // 19768

__host__ __device__ __forceinline__ int mlp_feat(int q, int g, int e) {
  return 32 * q + (e < 4 ? 4 * g + e : 16 + 4 * g + e - 4);
}

// packed network in device memory (bf16 bit patterns)
struct MlpPacked {
  uint16_t A1[8][64][4];      // layer 1, out tile r: W1[16r + l%16][4(l/16) + e]
  uint16_t A2[8][4][64][8];   // layer 2, out tile r, k-step q: W2[16r + l%16][feat(q, l/16, e)]
  uint16_t A3[4][64][8];      // layer 3, k-step q: W3[l%16][feat(q, l/16, e)] (rows 8..15: 0)
  uint16_t A3T[8][64][4];     // W3 transposed, tile r: W3[4(l/16)+e][16r + l%16] (e' >= 8: 0)
  uint16_t A2T[8][4][64][8];  // W2 transposed: W2[feat(q, l/16, e)][16r + l%16]
  float b1[128], b2[128], b3[16];
  float mu_in[16], isig_in[16], mu_out[8], sig_out[8];
};
// the part the inference kernels stage in LDS (A1, A2, A3 and the float block)
constexpr int kMlpInferBytes = sizeof(uint16_t) * (8 * 64 * 4 + 8 * 4 * 64 * 8 + 4 * 64 * 8);

// bf16 round to nearest even of an FP32 value (no NaN inputs on this path)
__host__ __device__ __forceinline__ uint16_t bf16_bits(float f) {
  uint32_t u;
  __builtin_memcpy(&u, &f, 4);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__host__ __device__ __forceinline__ float bf16_value(uint16_t b) {
  const uint32_t u = (uint32_t)b << 16;
  float f;
  __builtin_memcpy(&f, &u, 4);
  return f;
}

// element n of the packed arrays (A1 .. A2T flattened in declaration order) from the master
// parameters; the float block is copied separately
constexpr int kPackA1 = 8 * 64 * 4, kPackA2 = 8 * 4 * 64 * 8, kPackA3 = 4 * 64 * 8,
              kPackA3T = 8 * 64 * 4, kPackA2T = 8 * 4 * 64 * 8;
constexpr int kPackWeights = kPackA1 + kPackA2 + kPackA3 + kPackA3T + kPackA2T;
__host__ __device__ __forceinline__ uint16_t pack_weight(const float *P, int n) {
  if (n < kPackA1) {
    const int e = n % 4, l = (n / 4) % 64, r = n / 256;
    return bf16_bits(P[kMlpW1 + (16 * r + l % 16) * 16 + 4 * (l / 16) + e]);
  }
  n -= kPackA1;
  if (n < kPackA2) {
    const int e = n % 8, l = (n / 8) % 64, q = (n / 512) % 4, r = n / 2048;
    return bf16_bits(P[kMlpW2 + (16 * r + l % 16) * 128 + mlp_feat(q, l / 16, e)]);
  }
  n -= kPackA2;
  if (n < kPackA3) {
    const int e = n % 8, l = (n / 8) % 64, q = n / 512;
    return l % 16 < 8 ? bf16_bits(P[kMlpW3 + (l % 16) * 128 + mlp_feat(q, l / 16, e)]) : 0;
  }
  n -= kPackA3;
  if (n < kPackA3T) {
    const int e = n % 4, l = (n / 4) % 64, r = n / 256;
    const int o = 4 * (l / 16) + e;
    return o < 8 ? bf16_bits(P[kMlpW3 + o * 128 + 16 * r + l % 16]) : 0;
  }
  n -= kPackA3T;
  {
    const int e = n % 8, l = (n / 8) % 64, q = (n / 512) % 4, r = n / 2048;
    return bf16_bits(P[kMlpW2 + mlp_feat(q, l / 16, e) * 128 + 16 * r + l % 16]);
  }
}

}  // namespace ppk
