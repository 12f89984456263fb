// Training and evaluation kernels of the surrogate network (layout: mlp.hpp).
//
// One training step over a minibatch of B samples (B a multiple of 256):
//   k_fb      one wavefront per 64 samples: standardise, forward (MFMA), loss gradient,
//             backward to the pre-activation gradients (MFMA with the transposed weights);
//             stores the bf16 activations and gradients feature-major for the next kernel
//   k_wgrad   one block per 256 samples: weight gradients dW = dZ^T H over its samples
//             (MFMA, the samples are the k dimension) and bias sums, as FP32 partials
//   k_adam    sums the partials and applies Adam to the FP32 master parameters
//   k_pack    bf16 MFMA operand copies of the new weights
// The loss is the mean over samples and the 8 outputs of (y - (t - mu_out) / sig_out)^2.
#include <hip/hip_runtime.h>

#include "mlp_device.hpp"
#include "mlp_kernels.hpp"

namespace ppk {
namespace mlpk {

using namespace mlpd;

constexpr int kWaveSamples = 64;
constexpr int kChunk = 256;  // samples per k_wgrad block

// sample index of minibatch slot i of training step `step` (uniform with replacement)
__device__ __forceinline__ long long pick(unsigned long long seed, long long step, long long i,
                                          long long n) {
  unsigned long long z = seed + 0x9E3779B97F4A7C15ull * (unsigned long long)(step * 0x10001 + i + 1);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return (long long)(z % (unsigned long long)n);
}

// standardised bf16 features of the wave's samples into sX, raw targets into sT
__device__ __forceinline__ void stage_samples(const MlpView &w, const float *X, const float *T,
                                              long long idx, uint16_t *sX, float *sT) {
  const int lane = lane_id();
  const float *x = X + idx * kMlpIn;
#pragma unroll
  for (int i = 0; i < kMlpIn; ++i) sX[lane * 16 + i] = bf16_bits((x[i] - w.mu_in[i]) * w.isig_in[i]);
  if (T) {
    const float *t = T + idx * kMlpOut;
#pragma unroll
    for (int o = 0; o < kMlpOut; ++o) sT[lane * 8 + o] = (t[o] - w.mu_out[o]) / w.sig_out[o];
  }
}

// scratch of one minibatch, feature-major bf16: X 16 x B, H1 128 x B, H2 128 x B, dZ1 128 x B,
// dZ2 128 x B, dY 8 x B
struct FbScratch {
  uint16_t *X, *H1, *H2, *dZ1, *dZ2, *dY;
};
__host__ __device__ __forceinline__ FbScratch fb_scratch(uint16_t *base, long long B) {
  return {base, base + 16 * B, base + 144 * B, base + 272 * B, base + 400 * B, base + 528 * B};
}

__global__ void __launch_bounds__(256) k_fb(const MlpPacked *__restrict__ net,
                                            const float *__restrict__ X,
                                            const float *__restrict__ T, const long long *idx,
                                            long long n, unsigned long long seed, long long step,
                                            long long B, uint16_t *scratch, float inv_count) {
  __shared__ __attribute__((aligned(16))) uint16_t sXall[4][64 * 16];
  __shared__ float sTall[4][64 * 8];
  const MlpView w = view_of(net);
  const int wave = threadIdx.x >> 6, lane = lane_id(), g = lane >> 4, col = lane & 15;
  const long long s0 = ((long long)blockIdx.x * 4 + wave) * kWaveSamples;
  if (s0 >= B) return;
  uint16_t *sX = sXall[wave];
  float *sT = sTall[wave];
  const long long me = s0 + lane;
  const long long sample = idx ? idx[me] : pick(seed, step, me, n);
  stage_samples(w, X, T, sample, sX, sT);
  wave_lds_sync();
  FbScratch S = fb_scratch(scratch, B);
#pragma unroll
  for (int i = 0; i < 16; ++i) S.X[(long long)i * B + me] = sX[lane * 16 + i];
#pragma unroll 1
  for (int f = 0; f < 4; ++f) {
    const long long sc = s0 + 16 * f + col;  // the sample of this lane's column
    const v4s xb = *reinterpret_cast<const v4s *>(sX + (16 * f + col) * 16 + 4 * g);
    v8bf h1b[4], h2b[4];
    unsigned m1, m2;
    v4f y;
    const MlpView wf = opaque(w);
    mlp_forward16(wf, xb, h1b, h2b, m1, m2, y);
    // the bf16 activations as the forward pass used them (element e of pair q: feature
    // feat(q, g, e))
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const int h = mlp_feat(q, g, e);
        __bf16 a = h1b[q][e], b = h2b[q][e];
        uint16_t ua, ub;
        __builtin_memcpy(&ua, &a, 2);
        __builtin_memcpy(&ub, &b, 2);
        S.H1[(long long)h * B + sc] = ua;
        S.H2[(long long)h * B + sc] = ub;
      }
    // loss gradient (outputs 4g + j of sample col)
    v4s dyb;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int o = 4 * g + j;
      const float dy = o < 8 ? 2.f * (y[j] - sT[(16 * f + col) * 8 + o]) * inv_count : 0.f;
      dyb[j] = bf16s(dy);
      if (o < 8) S.dY[(long long)o * B + sc] = (uint16_t)dyb[j];
    }
    // dH2 = W3^T dY, dZ2 = dH2 * (z2 > 0), packed in pairs as the B operands of dH1
    v8bf db[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      v4f t[2];
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int r = 2 * q + h;
        v4f c = {0.f, 0.f, 0.f, 0.f};
        c = mfma_k16(ld_a16(wf.A3T, r, lane), dyb, c);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          t[h][j] = (m2 >> (4 * r + j)) & 1u ? c[j] : 0.f;
          S.dZ2[(long long)(16 * r + 4 * g + j) * B + sc] = bf16_bits(t[h][j]);
        }
      }
      db[q] = pack_pair(t[0], t[1]);
    }
    // dH1 = W2^T dZ2, dZ1 = dH1 * (z1 > 0)
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      v4f c = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int q = 0; q < 4; ++q) c = mfma_k32(ld_a32(wf.A2T, r * 4 + q, lane), db[q], c);
#pragma unroll
      for (int j = 0; j < 4; ++j)
        S.dZ1[(long long)(16 * r + 4 * g + j) * B + sc] =
            bf16_bits((m1 >> (4 * r + j)) & 1u ? c[j] : 0.f);
    }
  }
}

// one 16x16 tile of sum_s A[rowA + i][s] * Bm[rowB + n][s] over samples [s0, s0 + kChunk)
__device__ __forceinline__ v4f tile_gemm(const uint16_t *A, int rowA, int nrowA,
                                         const uint16_t *Bm, int rowB, long long B, long long s0) {
  const int lane = lane_id(), g = lane >> 4, col = lane & 15;
  v4f c = {0.f, 0.f, 0.f, 0.f};
  const int ra = rowA + col;
#pragma unroll 2
  for (int k = 0; k < kChunk; k += 32) {
    v8bf a, b;
    const long long off = s0 + k + 8 * g;
    if (ra < nrowA)
      a = *reinterpret_cast<const v8bf *>(A + (long long)ra * B + off);
    else
      a = v8bf{};
    b = *reinterpret_cast<const v8bf *>(Bm + (long long)(rowB + col) * B + off);
    c = mfma_k32(a, b, c);
  }
  return c;
}

__global__ void __launch_bounds__(256) k_wgrad(const uint16_t *scratch, long long B,
                                               float *partials) {
  const FbScratch S = fb_scratch(const_cast<uint16_t *>(scratch), B);
  const long long s0 = (long long)blockIdx.x * kChunk;
  float *P = partials + (long long)blockIdx.x * kMlpTrained;
  const int wave = threadIdx.x >> 6, lane = lane_id(), g = lane >> 4, col = lane & 15;
  // 80 tiles: dW2 64 (a, b), dW1 8 (a), dW3 8 (b); 20 per wave
  for (int t = wave; t < 80; t += 4) {
    if (t < 64) {
      const int a = t / 8, b = t % 8;
      const v4f c = tile_gemm(S.dZ2, 16 * a, 128, S.H1, 16 * b, B, s0);
#pragma unroll
      for (int j = 0; j < 4; ++j) P[kMlpW2 + (16 * a + 4 * g + j) * 128 + 16 * b + col] = c[j];
    } else if (t < 72) {
      const int a = t - 64;
      const v4f c = tile_gemm(S.dZ1, 16 * a, 128, S.X, 0, B, s0);
#pragma unroll
      for (int j = 0; j < 4; ++j) P[kMlpW1 + (16 * a + 4 * g + j) * 16 + col] = c[j];
    } else {
      const int b = t - 72;
      const v4f c = tile_gemm(S.dY, 0, 8, S.H2, 16 * b, B, s0);
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (4 * g + j < 8) P[kMlpW3 + (4 * g + j) * 128 + 16 * b + col] = c[j];
    }
  }
  // bias gradients: sums of the stored bf16 gradients over the chunk
  const int f = threadIdx.x;
  {
    const uint16_t *src = f < 128 ? S.dZ1 + (long long)f * B : S.dZ2 + (long long)(f - 128) * B;
    float acc = 0.f;
    for (int s = 0; s < kChunk; ++s) acc += bf16_value(src[s0 + s]);
    P[(f < 128 ? kMlpB1 + f : kMlpB2 + f - 128)] = acc;
  }
  if (f < 8) {
    float acc = 0.f;
    for (int s = 0; s < kChunk; ++s) acc += bf16_value(S.dY[(long long)f * B + s0 + s]);
    P[kMlpB3 + f] = acc;
  }
}

__global__ void k_adam(const float *partials, int nchunks, float *P, float *m, float *v,
                       float *grad_out, int step, float lr, float beta1, float beta2, float eps) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= kMlpTrained) return;
  float gsum = 0.f;
  for (int c = 0; c < nchunks; ++c) gsum += partials[(long long)c * kMlpTrained + i];
  if (grad_out) grad_out[i] = gsum;
  if (lr == 0.f) return;  // gradient only
  const float mi = beta1 * m[i] + (1.f - beta1) * gsum;
  const float vi = beta2 * v[i] + (1.f - beta2) * gsum * gsum;
  m[i] = mi;
  v[i] = vi;
  const float mh = mi / (1.f - __powf(beta1, (float)step));
  const float vh = vi / (1.f - __powf(beta2, (float)step));
  P[i] -= lr * mh / (sqrtf(vh) + eps);
}

__global__ void k_pack(const float *P, MlpPacked *net) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  uint16_t *W = &net->A1[0][0][0];
  if (n < kPackWeights) W[n] = pack_weight(P, n);
  if (n < 128) {
    net->b1[n] = P[kMlpB1 + n];
    net->b2[n] = P[kMlpB2 + n];
  }
  if (n < 16) {
    net->b3[n] = n < 8 ? P[kMlpB3 + n] : 0.f;
    net->mu_in[n] = P[kMlpMuIn + n];
    net->isig_in[n] = P[kMlpIsigIn + n];
  }
  if (n < 8) {
    net->mu_out[n] = P[kMlpMuOut + n];
    net->sig_out[n] = P[kMlpSigOut + n];
  }
}

// inference on n samples (features X[n][16]) -> Y[n][8] (standardised outputs) and, with
// targets, per-block partial sums of squared errors per output
__global__ void __launch_bounds__(256) k_infer(const MlpPacked *__restrict__ net,
                                               const float *__restrict__ X,
                                               const float *__restrict__ T, long long n,
                                               float *Y, double *sq_parts) {
  __shared__ __attribute__((aligned(16))) uint16_t sXall[4][64 * 16];
  __shared__ float sTall[4][64 * 8];
  __shared__ float sYall[4][64 * 8];
  __shared__ double red[256];
  const MlpView w = view_of(net);
  const int wave = threadIdx.x >> 6, lane = lane_id();
  const long long me = ((long long)blockIdx.x * 4 + wave) * kWaveSamples + lane;
  uint16_t *sX = sXall[wave];
  float *sT = sTall[wave], *sY = sYall[wave];
  stage_samples(w, X, T, me < n ? me : n - 1, sX, sT);
  wave_lds_sync();
  mlp_infer_wave(w, sX, sY);
  double sq[8];
#pragma unroll
  for (int o = 0; o < 8; ++o) {
    const float y = sY[lane * 8 + o];
    if (me < n && Y) Y[me * 8 + o] = y;
    const double d = T && me < n ? (double)y - (double)sT[lane * 8 + o] : 0.0;
    sq[o] = d * d;
  }
  if (!sq_parts) return;
#pragma unroll 1
  for (int o = 0; o < 8; ++o) {
    red[threadIdx.x] = sq[o];
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
      if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
      __syncthreads();
    }
    if (threadIdx.x == 0) sq_parts[(long long)blockIdx.x * 8 + o] = red[0];
    __syncthreads();
  }
}

// per-block sums and sums of squares of the 16 features and 8 targets (standardisation)
__global__ void __launch_bounds__(256) k_moments(const float *__restrict__ X,
                                                 const float *__restrict__ T, long long n,
                                                 double *parts) {
  __shared__ double red[256];
  const long long per = (n + gridDim.x - 1) / gridDim.x;
  const long long lo = (long long)blockIdx.x * per, hi = lo + per < n ? lo + per : n;
#pragma unroll 1
  for (int c = 0; c < 24; ++c) {
    double s = 0.0, s2 = 0.0;
    for (long long i = lo + threadIdx.x; i < hi; i += 256) {
      const double v = c < 16 ? X[i * 16 + c] : T[i * 8 + c - 16];
      s += v;
      s2 += v * v;
    }
    for (int k = 0; k < 2; ++k) {
      red[threadIdx.x] = k ? s2 : s;
      __syncthreads();
      for (int st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
        __syncthreads();
      }
      if (threadIdx.x == 0) parts[((long long)blockIdx.x * 24 + c) * 2 + k] = red[0];
      __syncthreads();
    }
  }
}

}  // namespace mlpk

// ---- launchers (mlp_kernels.hpp) -----------------------------------------------------------
namespace mlp {

size_t fb_scratch_elems(long long B) { return (size_t)536 * B; }

hipError_t train_step(const MlpPacked *net, float *P, float *m, float *v, const float *X,
                      const float *T, const long long *idx, long long n, unsigned long long seed,
                      long long step, long long B, uint16_t *scratch, float *partials,
                      float *grad_out, float lr, hipStream_t s) {
  using namespace mlpk;
  if (B % kChunk) return hipErrorInvalidValue;
  const float inv_count = 1.f / (float)(B * kMlpOut);
  hipLaunchKernelGGL(k_fb, dim3((unsigned)(B / 256)), dim3(256), 0, s, net, X, T, idx, n, seed,
                     step, B, scratch, inv_count);
  hipLaunchKernelGGL(k_wgrad, dim3((unsigned)(B / kChunk)), dim3(256), 0, s, scratch, B, partials);
  hipLaunchKernelGGL(k_adam, dim3((kMlpTrained + 255) / 256), dim3(256), 0, s, partials,
                     (int)(B / kChunk), P, m, v, grad_out, (int)(step + 1), lr, 0.9f, 0.999f,
                     1e-8f);
  if (lr != 0.f)
    hipLaunchKernelGGL(k_pack, dim3((kPackWeights + 255) / 256), dim3(256), 0, s, P,
                       const_cast<MlpPacked *>(net));
  return hipGetLastError();
}

hipError_t pack(const float *P, MlpPacked *net, hipStream_t s) {
  hipLaunchKernelGGL(mlpk::k_pack, dim3((kPackWeights + 255) / 256), dim3(256), 0, s, P, net);
  return hipGetLastError();
}

int infer_blocks(long long n) { return (int)((n + 255) / 256); }

hipError_t infer(const MlpPacked *net, const float *X, const float *T, long long n, float *Y,
                 double *sq_parts, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(mlpk::k_infer, dim3((unsigned)infer_blocks(n)), dim3(256), 0, s, net, X, T,
                     n, Y, sq_parts);
  return hipGetLastError();
}

hipError_t moments(const float *X, const float *T, long long n, double *parts, int nblocks,
                   hipStream_t s) {
  hipLaunchKernelGGL(mlpk::k_moments, dim3(nblocks), dim3(256), 0, s, X, T, n, parts);
  return hipGetLastError();
}

}  // namespace mlp
}  // namespace ppk
