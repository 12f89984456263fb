// Wave-level MLP on the CDNA4 matrix cores: one wavefront evaluates 64 samples as four blocks
// of 16, each layer a row of 16x16 MFMA tiles (bf16 operands, FP32 accumulators).  Layout and
// weight packing: mlp.hpp.  Device code only.
#pragma once

#include <hip/hip_runtime.h>

#include "mlp.hpp"

namespace ppk {
namespace mlpd {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef short v2s __attribute__((ext_vector_type(2)));
typedef unsigned v4u __attribute__((ext_vector_type(4)));
typedef float v2f __attribute__((ext_vector_type(2)));
typedef __bf16 v2bf __attribute__((ext_vector_type(2)));

__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }

// make LDS writes of this wavefront visible to its other lanes
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ v4f mfma_k16(v4s a, v4s b, v4f c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ v4f mfma_k32(v8bf a, v8bf b, v4f c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// bf16 round to nearest even in hardware (v_cvt_pk_bf16_f32; the same rounding as bf16_bits)
__device__ __forceinline__ __bf16 bf16b(float x) { return static_cast<__bf16>(x); }
__device__ __forceinline__ short bf16s(float x) { return __builtin_bit_cast(short, bf16b(x)); }
__device__ __forceinline__ float relu(float x) { return x > 0.f ? x : 0.f; }

// the pieces of a packed network the wave functions read (LDS or global)
struct MlpView {
  const uint16_t *A1, *A2, *A3, *A3T, *A2T;
  const float *b1, *b2, *b3, *mu_in, *isig_in, *mu_out, *sig_out;
};
__device__ __forceinline__ MlpView view_of(const MlpPacked *w) {
  return {&w->A1[0][0][0], &w->A2[0][0][0][0], &w->A3[0][0][0], &w->A3T[0][0][0],
          &w->A2T[0][0][0][0], w->b1, w->b2, w->b3, w->mu_in, w->isig_in, w->mu_out,
          w->sig_out};
}

// Hide the weight pointers from loop-invariant code motion: every 16-sample block reloads its
// operands (cache / LDS hits) instead of the compiler holding all ~40 KB of weight operands of
// a wave in registers across the loop.
// The pointers move by an opaque zero rather than being hidden themselves, so the compiler
// still sees where they point (LDS weights stay ds_read, not flat loads).
__device__ __forceinline__ MlpView opaque(MlpView w) {
  int z = 0;
  asm volatile("" : "+s"(z));
  w.A1 += z;
  w.A2 += z;
  w.A3 += z;
  if (w.A3T) w.A3T += z;
  if (w.A2T) w.A2T += z;
  w.b1 += z;
  w.b2 += z;
  w.b3 += z;
  return w;
}

__device__ __forceinline__ v4s ld_a16(const uint16_t *A, int tile, int lane) {
  return *reinterpret_cast<const v4s *>(A + ((size_t)tile * 64 + lane) * 4);
}
__device__ __forceinline__ v8bf ld_a32(const uint16_t *A, int tile, int lane) {
  return *reinterpret_cast<const v8bf *>(A + ((size_t)tile * 64 + lane) * 8);
}

// two floats rounded to bf16 (nearest even) in one 32-bit word, a in the low half (one
// v_cvt_pk_bf16_f32)
__device__ __forceinline__ unsigned bf16x2(float a, float b) {
  return __builtin_bit_cast(unsigned, __builtin_convertvector((v2f){a, b}, v2bf));
}

// bf16 B operand of a 16x16x32 step q from two accumulator tiles (features 32q .. 32q+31)
__device__ __forceinline__ v8bf pack_pair(v4f lo, v4f hi) {
  const v4u r = {bf16x2(lo[0], lo[1]), bf16x2(lo[2], lo[3]), bf16x2(hi[0], hi[1]),
                 bf16x2(hi[2], hi[3])};
  return __builtin_bit_cast(v8bf, r);
}

// ReLU of the two bf16 values of a word, as 16-bit integers (one v_pk_max_i16)
__device__ __forceinline__ unsigned relu_bf16x2(unsigned w) {
  const v2s zero = {0, 0};
  return __builtin_bit_cast(unsigned, __builtin_elementwise_max(__builtin_bit_cast(v2s, w), zero));
}

// ReLU of two accumulator tiles, packed like pack_pair.  The tiles are rounded to bf16 first
// and the ReLU taken on the bf16 bit patterns as 16-bit integers (one v_pk_max_i16 per pair):
// a negative value or -0 has the sign bit set, so its pattern is a negative integer and the
// maximum with 0 is +0; rounding to nearest even keeps the sign, so the result equals
// bf16(max(x, 0)) for every non-NaN x.
__device__ __forceinline__ v8bf pack_relu(v4f lo, v4f hi) {
  const v4u r = {relu_bf16x2(bf16x2(lo[0], lo[1])), relu_bf16x2(bf16x2(lo[2], lo[3])),
                 relu_bf16x2(bf16x2(hi[0], hi[1])), relu_bf16x2(bf16x2(hi[2], hi[3]))};
  return __builtin_bit_cast(v8bf, r);
}

// Forward pass of one block of 16 samples.  xb: B operand of layer 1 (lane l: the bf16
// features 4(l/16) .. 4(l/16)+3 of sample l%16).  Each hidden tile is rounded to bf16 as soon
// as it is complete (tile r: features 16r + 4(l/16) + j of sample l%16), packed in pairs into
// the next layer's B operands h1b / h2b; the masks hold the signs of the pre-activations (bit
// 4r + j).  y: outputs 4(l/16) + j (rows >= 8 are 0).
__device__ __forceinline__ void mlp_forward16(const MlpView &w, v4s xb, v8bf h1b[4], v8bf h2b[4],
                                              unsigned &mask1, unsigned &mask2, v4f &y) {
  const int lane = lane_id(), g = lane >> 4;
  mask1 = mask2 = 0u;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    v4f t[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int r = 2 * q + h;
      v4f c;
#pragma unroll
      for (int j = 0; j < 4; ++j) c[j] = w.b1[16 * r + 4 * g + j];
      c = mfma_k16(ld_a16(w.A1, r, lane), xb, c);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        mask1 |= (c[j] > 0.f ? 1u : 0u) << (4 * r + j);
        t[h][j] = relu(c[j]);
      }
    }
    h1b[q] = pack_pair(t[0], t[1]);
  }
#pragma unroll
  for (int q2 = 0; q2 < 4; ++q2) {
    v4f t[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int r = 2 * q2 + h;
      v4f c;
#pragma unroll
      for (int j = 0; j < 4; ++j) c[j] = w.b2[16 * r + 4 * g + j];
#pragma unroll
      for (int q = 0; q < 4; ++q) c = mfma_k32(ld_a32(w.A2, r * 4 + q, lane), h1b[q], c);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        mask2 |= (c[j] > 0.f ? 1u : 0u) << (4 * r + j);
        t[h][j] = relu(c[j]);
      }
    }
    h2b[q2] = pack_pair(t[0], t[1]);
  }
  v4f c;
#pragma unroll
  for (int j = 0; j < 4; ++j) c[j] = w.b3[4 * g + j];
#pragma unroll
  for (int q = 0; q < 4; ++q) c = mfma_k32(ld_a32(w.A3, q, lane), h2b[q], c);
  y = c;
}

// Inference of the 64 samples of a wave: sX holds their standardised bf16 features
// ([sample][16], this wave's slice of LDS), sY receives the 8 outputs per sample
// ([sample][8] floats).  Every lane of the wave must call it.  NB blocks of 16 samples share
// each weight operand load (NB = 4: every A operand is read once per wave).
template <int NB>
__device__ __forceinline__ void mlp_infer_blocks(const MlpView &w, const uint16_t *sX, float *sY,
                                                 int f0) {
  const int lane = lane_id(), g = lane >> 4, col = lane & 15;
  v4s xb[NB];
#pragma unroll
  for (int f = 0; f < NB; ++f)
    xb[f] = *reinterpret_cast<const v4s *>(sX + (16 * (f0 + f) + col) * 16 + 4 * g);
  v8bf h1b[NB][4], h2b[NB][4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    v4f t[NB][2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int r = 2 * q + h;
      v4f bias;
#pragma unroll
      for (int j = 0; j < 4; ++j) bias[j] = w.b1[16 * r + 4 * g + j];
      const v4s a = ld_a16(w.A1, r, lane);
#pragma unroll
      for (int f = 0; f < NB; ++f) t[f][h] = mfma_k16(a, xb[f], bias);
    }
#pragma unroll
    for (int f = 0; f < NB; ++f) h1b[f][q] = pack_relu(t[f][0], t[f][1]);
  }
#pragma unroll
  for (int q2 = 0; q2 < 4; ++q2) {
    v4f t[NB][2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int r = 2 * q2 + h;
      v4f acc[NB];
#pragma unroll
      for (int f = 0; f < NB; ++f)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[f][j] = w.b2[16 * r + 4 * g + j];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const v8bf a = ld_a32(w.A2, r * 4 + q, lane);
#pragma unroll
        for (int f = 0; f < NB; ++f) acc[f] = mfma_k32(a, h1b[f][q], acc[f]);
      }
#pragma unroll
      for (int f = 0; f < NB; ++f) t[f][h] = acc[f];
    }
#pragma unroll
    for (int f = 0; f < NB; ++f) h2b[f][q2] = pack_relu(t[f][0], t[f][1]);
  }
  v4f y[NB];
#pragma unroll
  for (int f = 0; f < NB; ++f)
#pragma unroll
    for (int j = 0; j < 4; ++j) y[f][j] = w.b3[4 * g + j];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const v8bf a = ld_a32(w.A3, q, lane);
#pragma unroll
    for (int f = 0; f < NB; ++f) y[f] = mfma_k32(a, h2b[f][q], y[f]);
  }
  if (g < 2) {
#pragma unroll
    for (int f = 0; f < NB; ++f)
#pragma unroll
      for (int j = 0; j < 4; ++j) sY[(16 * (f0 + f) + col) * 8 + 4 * g + j] = y[f][j];
  }
}

template <int NB = 4>
__device__ __forceinline__ void mlp_infer_wave(const MlpView &w, const uint16_t *sX, float *sY) {
#pragma unroll 1
  for (int f0 = 0; f0 < 4; f0 += NB) mlp_infer_blocks<NB>(opaque(w), sX, sY, f0);
  wave_lds_sync();
}

}  // namespace mlpd
}  // namespace ppk
