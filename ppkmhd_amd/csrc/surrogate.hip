// Surrogate form of the face solves, and the face-sample dump its network is trained on.
//
//   k_face_sur<DIM>     the face fluxes of every direction with the learned flux where the gate
//                       and the admissibility check allow (per face).  A block of 256 threads
//                       stages the network (40 KB of bf16 weights) in LDS once and then walks
//                       bricks of the face sub-box, direction after direction; each wavefront
//                       runs the network on the matrix cores for its 64 faces when any of them
//                       passed the gate, and marks the faces that did not or whose flux was
//                       refused
//   k_face_fallback     exact HLLD for the marked faces (every direction in one launch)
//   k_dump<DIM>         features and training targets of every stride-th face (all directions)
//
// The face states are those of the plain fused step (face_states2/3), so a face that falls
// back gets the plain step's flux bit for bit.  Formulas of the surrogate: surrogate_body.hpp.
#include <hip/hip_runtime.h>

#include "fused_body.hpp"
#include "mlp_device.hpp"
#include "surrogate_body.hpp"

namespace ppk {
namespace PPK_NS {
namespace surk {

using namespace fusedk;
using namespace mlpd;

constexpr int kBlock = 256;
#ifndef PPK_SUR_NB
#define PPK_SUR_NB 2
#endif
constexpr int kNB = PPK_SUR_NB;  // 16-face blocks per weight-operand load
constexpr int kFloatBlock = 128 + 128 + 16 + 16 + 16 + 8 + 8;  // b1 b2 b3 mu_in isig_in mu_out sig_out

__device__ __forceinline__ bool active(const double *ctrl) { return ctrl[CTRL_ACTIVE] != 0.0; }

template <int DIM, int D, class R, class T>
PPK_HD void states(const R *rec, const DevParams &p, long long c, St<T> &ql, St<T> &qr) {
  if constexpr (DIM == 3)
    face_states3<D>(rec, p, c, ql, qr);
  else
    face_states2<D>(rec, p, c, ql, qr);
}
template <int DIM, int D, class R>
PPK_HD void store(R *flux, const DevParams &p, long long c, const St<R> &f) {
  if constexpr (DIM == 3)
    store_flux3<D>(flux, p, c, f);
  else
    store_flux2<D>(flux, p, c, f);
}

// the faces of direction D of the bricks this block walks (see k_face_sur); w: the network in
// LDS, sX: this wave's LDS slice; counts of faces by the network / in total accumulate
template <int DIM, int D, class R>
__device__ __forceinline__ void faces_dir(const R *__restrict__ rec, R *__restrict__ flux,
                                          const DevParams &p, const Brick &bk,
                                          const SurrogateCfg &cfg, const MlpView &w,
                                          uint16_t *sX, unsigned long long &n_net,
                                          unsigned long long &n_all) {
  const int lane = lane_id();
  float *sY = reinterpret_cast<float *>(sX);
  const float gamma0 = float(p.gamma0);
  const long long nb = bk.nblocks;
#pragma unroll 1
  for (long long L = blockIdx.x; L < nb; L += gridDim.x) {
    const long long c = sub_brick_cell(p, bk, (int)L, threadIdx.x);
    // the face states, the LLF pieces, the features and the check in float: their rounding
    // is far below the network's own error; faces left to HLLD are recomputed in R there
    St<float> a, b;
    St<R> F;
    bool gate = false;
    if (c >= 0) {
      states<DIM, D>(rec, p, c, a, b);
      gate = smooth_face(a, b, cfg.gate_rho, cfg.gate_p);
    }
    bool use = false;
    if (__ballot(gate) != 0ull) {  // the network runs for the whole wavefront (wave-uniform)
      St<float> Fl;
      float s = 1.f, rb = 1.f;
      float x[kMlpIn];
      if (gate) {
        St<float> ul, ur, fl, fr;
        llf_parts(a, b, ul, ur, fl, fr, s, Fl, p);
        face_features(a, b, s, gamma0, x, rb);
      } else {
#pragma unroll
        for (int i = 0; i < kMlpIn; ++i) x[i] = w.mu_in[i];
      }
#pragma unroll
      for (int i = 0; i < kMlpIn; ++i)
        sX[lane * 16 + i] = __builtin_bit_cast(uint16_t, bf16b((x[i] - w.mu_in[i]) * w.isig_in[i]));
      wave_lds_sync();
      mlp_infer_wave<kNB>(w, sX, sY);
      if (gate) {
        float sc[kMlpOut];
        out_scales(rb, s, sc);
        St<float> Fs, ul, ur, fl, fr;
#pragma unroll
        for (int o = 0; o < kMlpOut; ++o)
          Fs[o] = Fl[o] + sc[o] * (w.mu_out[o] + w.sig_out[o] * sY[lane * 8 + o]);
        mhd_flux(a, ul, fl, p);
        mhd_flux(b, ur, fr, p);
        use = admissible(ul, ur, fl, fr, Fs, s, gamma0);
        if (use) {
#pragma unroll
          for (int o = 0; o < 8; ++o) F[o] = R(Fs[o]);
        }
      }
    }
    // the network's flux, or a mark for the HLLD pass (which keeps this kernel's registers
    // free of the HLLD solver)
    if (c >= 0) {
      if (use) store<DIM, D>(flux, p, c, F);
      cfg.flags[(long long)D * p.vstride + c] = use ? 0 : 1;
    }
    n_net += __popcll(__ballot(c >= 0 && use));
    n_all += __popcll(__ballot(c >= 0));
  }
}

// The learned face fluxes of every direction in one launch: a block stages the network in
// LDS once and walks its bricks for direction x, then y (then z); see the file comment.
template <int DIM, class R>
__global__ void __launch_bounds__(kBlock) k_face_sur(const R *__restrict__ rec,
                                                     R *__restrict__ flux, DevParams p, Brick bk,
                                                     const double *ctrl, SurrogateCfg cfg) {
  if (!active(ctrl)) return;
  __shared__ __attribute__((aligned(16))) uint16_t sW[kPackA1 + kPackA2 + kPackA3];
  __shared__ float sFl[kFloatBlock];
  // per wave: the bf16 features of its 64 faces, then (same bytes, 32 per face) the outputs:
  // a pass of the network reads the features of its 16-face blocks before it writes their
  // outputs, and later passes read only later blocks
  __shared__ __attribute__((aligned(16))) uint16_t sXall[4][64 * 16];
  const MlpPacked *net = static_cast<const MlpPacked *>(cfg.net);
  {  // stage the inference weights: A1, A2, A3 are the first kMlpInferBytes of the net
    const uint4 *src = reinterpret_cast<const uint4 *>(net);
    uint4 *dst = reinterpret_cast<uint4 *>(sW);
    for (int n = threadIdx.x; n < kMlpInferBytes / 16; n += kBlock) dst[n] = src[n];
    const float *fsrc = net->b1;  // b1 .. sig_out are contiguous
    for (int n = threadIdx.x; n < kFloatBlock; n += kBlock) sFl[n] = fsrc[n];
  }
  __syncthreads();
  MlpView w;
  w.A1 = sW;
  w.A2 = sW + kPackA1;
  w.A3 = sW + kPackA1 + kPackA2;
  w.A3T = w.A2T = nullptr;
  w.b1 = sFl;
  w.b2 = sFl + 128;
  w.b3 = sFl + 256;
  w.mu_in = sFl + 272;
  w.isig_in = sFl + 288;
  w.mu_out = sFl + 304;
  w.sig_out = sFl + 312;
  const int wave = threadIdx.x >> 6, lane = lane_id();
  uint16_t *sX = sXall[wave];
  unsigned long long n_net = 0, n_all = 0;  // faces of this wave by the network, in total
  faces_dir<DIM, 0, R>(rec, flux, p, bk, cfg, w, sX, n_net, n_all);
  faces_dir<DIM, 1, R>(rec, flux, p, bk, cfg, w, sX, n_net, n_all);
  if constexpr (DIM == 3) faces_dir<DIM, 2, R>(rec, flux, p, bk, cfg, w, sX, n_net, n_all);
  // per-block counts (no same-address atomics): this block's slots, in stream order
  if (cfg.counters) {
    __shared__ unsigned long long cnt[2][kBlock / 64];
    if (lane == 0) {
      cnt[0][wave] = n_net;
      cnt[1][wave] = n_all;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      unsigned long long a = 0, b = 0;
      for (int i = 0; i < kBlock / 64; ++i) {
        a += cnt[0][i];
        b += cnt[1][i];
      }
      cfg.counters[2 * blockIdx.x] += a;
      cfg.counters[2 * blockIdx.x + 1] += b - a;
    }
  }
}

// exact HLLD for the faces of direction D that the surrogate pass marked
template <int DIM, int D, class R>
__device__ __forceinline__ void fallback_dir(const R *__restrict__ rec, R *__restrict__ flux,
                                             const DevParams &p, const Brick &bk, int L,
                                             const unsigned char *flags) {
  const long long c = sub_brick_cell(p, bk, L, threadIdx.x);
  if (c < 0 || !flags[(long long)D * p.vstride + c]) return;
  St<R> ql, qr, F;
  states<DIM, D>(rec, p, c, ql, qr);
  riemann_hlld(ql, qr, F, p);
  store<DIM, D>(flux, p, c, F);
}
// every direction in one launch: block B handles brick B / DIM of direction B % DIM
template <int DIM, class R>
__global__ void __launch_bounds__(kBlock) k_face_fallback(const R *__restrict__ rec,
                                                          R *__restrict__ flux, DevParams p,
                                                          Brick bk, const double *ctrl,
                                                          const unsigned char *flags) {
  if (!active(ctrl)) return;
  const int L = (int)(blockIdx.x / DIM), d = (int)(blockIdx.x % DIM);
  if (d == 0)
    fallback_dir<DIM, 0, R>(rec, flux, p, bk, L, flags);
  else if (d == 1)
    fallback_dir<DIM, 1, R>(rec, flux, p, bk, L, flags);
  else if constexpr (DIM == 3)
    fallback_dir<DIM, 2, R>(rec, flux, p, bk, L, flags);
}

// one thread per selected face: slot s is face t = s * stride + phase of the enumeration
// (direction, then the sub-box cells, i fastest)
template <int DIM, class R>
__global__ void __launch_bounds__(kBlock) k_dump(const R *__restrict__ rec, DevParams p,
                                                 const double *ctrl, float *X, float *T,
                                                 long long nslots, long long stride,
                                                 long long phase) {
  if (!active(ctrl)) return;
  const long long sl = (long long)blockIdx.x * kBlock + threadIdx.x;
  if (sl >= nslots) return;
  const long long t = sl * stride + phase;
  const long long nxs = p.nx + 1, nys = p.ny + 1, nzs = DIM == 3 ? p.nz + 1 : 1;
  const long long nsub = nxs * nys * nzs;
  const int D = (int)(t / nsub);
  const long long r = t % nsub;
  const int i = p.gw + (int)(r % nxs), j = p.gw + (int)((r / nxs) % nys);
  const int k = DIM == 3 ? p.gw + (int)(r / (nxs * nys)) : 0;
  const long long c = cidx(p, i, j, k);
  St<R> ql, qr;
  if (D == 0)
    states<DIM, 0>(rec, p, c, ql, qr);
  else if (D == 1)
    states<DIM, 1>(rec, p, c, ql, qr);
  else
    states<DIM, DIM == 3 ? 2 : 1>(rec, p, c, ql, qr);
  St<R> fh;
  {
    St<R> a = ql, b = qr;
    riemann_hlld(a, b, fh, p);
  }
  St<R> ul, ur, fl, fr, F;
  R s;
  llf_parts(ql, qr, ul, ur, fl, fr, s, F, p);
  float x[kMlpIn], tg[kMlpOut];
  R rb, sc[kMlpOut];
  face_features(ql, qr, s, R(p.gamma0), x, rb);
  out_scales(rb, s, sc);
  face_target(fh, F, sc, tg);
#pragma unroll
  for (int n = 0; n < kMlpIn; ++n) X[sl * kMlpIn + n] = x[n];
#pragma unroll
  for (int n = 0; n < kMlpOut; ++n) T[sl * kMlpOut + n] = tg[n];
}

template <class R>
hipError_t faces_impl(const DevParams &p, const R *rec, R *flux, const double *ctrl,
                      const SurrogateCfg &cfg, hipStream_t s, StageTimer *tm) {
  const Brick bk = sub_bricks(p);
  const long long nb = bk.count();
  const dim3 g((unsigned)(nb < kSurrogateBlocks ? nb : kSurrogateBlocks)), b(kBlock);
  const dim3 gf((unsigned)(p.dim * nb));
  const unsigned char *fl = cfg.flags;
  if (tm) tm->mark(ST_SUR_FACES, true);
  if (p.dim == 3)
    hipLaunchKernelGGL((k_face_sur<3, R>), g, b, 0, s, rec, flux, p, bk, ctrl, cfg);
  else
    hipLaunchKernelGGL((k_face_sur<2, R>), g, b, 0, s, rec, flux, p, bk, ctrl, cfg);
  if (tm) tm->mark(ST_SUR_FACES, false);
  if (tm) tm->mark(ST_SUR_FALLBACK, true);
  if (p.dim == 3)
    hipLaunchKernelGGL((k_face_fallback<3, R>), gf, b, 0, s, rec, flux, p, bk, ctrl, fl);
  else
    hipLaunchKernelGGL((k_face_fallback<2, R>), gf, b, 0, s, rec, flux, p, bk, ctrl, fl);
  if (tm) tm->mark(ST_SUR_FALLBACK, false);
  return hipGetLastError();
}

}  // namespace surk

long long surrogate_dump_slots(const DevParams &p, long long stride, long long phase) {
  const long long nsub = (long long)(p.nx + 1) * (p.ny + 1) * (p.dim == 3 ? p.nz + 1 : 1);
  const long long faces = p.dim * nsub;
  if (stride < 1 || phase < 0 || phase >= stride || phase >= faces) return 0;
  return (faces - phase + stride - 1) / stride;
}

hipError_t surrogate_faces(const DevParams &p, const void *rec, void *flux, const double *ctrl,
                           const SurrogateCfg &cfg, hipStream_t s, StageTimer *tm) {
  if (p.bulk32)
    return surk::faces_impl<float>(p, static_cast<const float *>(rec), static_cast<float *>(flux),
                                   ctrl, cfg, s, tm);
  return surk::faces_impl<double>(p, static_cast<const double *>(rec),
                                  static_cast<double *>(flux), ctrl, cfg, s, tm);
}

hipError_t surrogate_dump(const DevParams &p, const void *rec, const double *ctrl, float *X,
                          float *T, long long stride, long long phase, hipStream_t s) {
  const long long n = surrogate_dump_slots(p, stride, phase);
  if (n <= 0) return hipSuccess;
  const dim3 g((unsigned)((n + surk::kBlock - 1) / surk::kBlock)), b(surk::kBlock);
  if (p.dim == 3) {
    if (p.bulk32)
      hipLaunchKernelGGL((surk::k_dump<3, float>), g, b, 0, s, static_cast<const float *>(rec), p,
                         ctrl, X, T, n, stride, phase);
    else
      hipLaunchKernelGGL((surk::k_dump<3, double>), g, b, 0, s, static_cast<const double *>(rec),
                         p, ctrl, X, T, n, stride, phase);
  } else {
    if (p.bulk32)
      hipLaunchKernelGGL((surk::k_dump<2, float>), g, b, 0, s, static_cast<const float *>(rec), p,
                         ctrl, X, T, n, stride, phase);
    else
      hipLaunchKernelGGL((surk::k_dump<2, double>), g, b, 0, s, static_cast<const double *>(rec),
                         p, ctrl, X, T, n, stride, phase);
  }
  return hipGetLastError();
}

}  // namespace PPK_NS
}  // namespace ppk
