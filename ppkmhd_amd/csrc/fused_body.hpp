// Per-thread bodies of the fused MHD step (host+device): records, the LDS-tiled half step, the
// face and edge solves, the update with CT and the CFL term, and the block-to-cells mappings.
// Included by fused.hip (one grid per launch) and batch.hip (many grids per launch); the host
// emulation of the kernels includes it too.  Arithmetic references are given where each body
// is defined.
#pragma once

#include <hip/hip_runtime.h>

#include "kernels.hpp"
#include "trace.hpp"

namespace ppk {
namespace PPK_NS {
namespace fusedk {

// record fields, 3D: centre at t^{n+1/2}, halved slopes, left face fields at t^{n+1/2} and
// their transverse slopes
enum {
  R3_R = 0, R3_P, R3_U, R3_V, R3_W, R3_A, R3_B, R3_C,
  R3_DRX, R3_DPX, R3_DUX, R3_DVX, R3_DWX, R3_DBX, R3_DCX,
  R3_DRY, R3_DPY, R3_DUY, R3_DVY, R3_DWY, R3_DAY, R3_DCY,
  R3_DRZ, R3_DPZ, R3_DUZ, R3_DVZ, R3_DWZ, R3_DAZ, R3_DBZ,
  R3_AL, R3_BL, R3_CL, R3_DALY, R3_DALZ, R3_DBLX, R3_DBLZ, R3_DCLX, R3_DCLY,
  R3_N
};
static_assert(R3_N == kRec3, "3D record size");
// record fields, 2D
enum {
  R2_R = 0, R2_P, R2_U, R2_V, R2_W, R2_A, R2_B, R2_C,
  R2_DRX, R2_DPX, R2_DUX, R2_DVX, R2_DWX, R2_DBX, R2_DCX,
  R2_DRY, R2_DPY, R2_DUY, R2_DVY, R2_DWY, R2_DAY, R2_DCY,
  R2_AL, R2_BL, R2_DALY, R2_DBLX,
  R2_N
};
static_assert(R2_N == kRec2, "2D record size");

PPK_HD long long cidx(const DevParams &p, int i, int j, int k) {
  return ((long long)k * p.jsize + j) * p.isize + i;
}

// Blocks are dealt to the 8 XCDs round-robin (block b runs on XCD b % 8).  Renumber them so
// that each XCD works through one contiguous range of tiles: neighbouring tiles then share
// that XCD's L2.
PPK_HD int xcd_remap(int b, int nblocks) {
  const int q = nblocks / 8, r = nblocks % 8, x = b % 8;
  return x * q + (x < r ? x : r) + b / 8;
}

template <class R>
PPK_HD void store_rec3(R *__restrict__ rec, long long vs, long long t, const Half3D<R> &h) {
  R *o = rec + t;
  o[R3_R * vs] = h.r;     o[R3_P * vs] = h.pp;    o[R3_U * vs] = h.u;     o[R3_V * vs] = h.v;
  o[R3_W * vs] = h.w;     o[R3_A * vs] = h.A;     o[R3_B * vs] = h.B;     o[R3_C * vs] = h.C;
  o[R3_DRX * vs] = h.drx; o[R3_DPX * vs] = h.dpx; o[R3_DUX * vs] = h.dux; o[R3_DVX * vs] = h.dvx;
  o[R3_DWX * vs] = h.dwx; o[R3_DBX * vs] = h.dBx; o[R3_DCX * vs] = h.dCx;
  o[R3_DRY * vs] = h.dry; o[R3_DPY * vs] = h.dpy; o[R3_DUY * vs] = h.duy; o[R3_DVY * vs] = h.dvy;
  o[R3_DWY * vs] = h.dwy; o[R3_DAY * vs] = h.dAy; o[R3_DCY * vs] = h.dCy;
  o[R3_DRZ * vs] = h.drz; o[R3_DPZ * vs] = h.dpz; o[R3_DUZ * vs] = h.duz; o[R3_DVZ * vs] = h.dvz;
  o[R3_DWZ * vs] = h.dwz; o[R3_DAZ * vs] = h.dAz; o[R3_DBZ * vs] = h.dBz;
  o[R3_AL * vs] = h.AL;   o[R3_BL * vs] = h.BL;   o[R3_CL * vs] = h.CL;
  o[R3_DALY * vs] = h.dALy; o[R3_DALZ * vs] = h.dALz; o[R3_DBLX * vs] = h.dBLx;
  o[R3_DBLZ * vs] = h.dBLz; o[R3_DCLX * vs] = h.dCLx; o[R3_DCLY * vs] = h.dCLy;
}

// left-side fields of the record at t; the right-side ones are set by the caller from the +1
// neighbours where a state needs them
template <class R, class T>
PPK_HD void load_rec3(const R *__restrict__ rec, long long vs, long long t, Half3D<T> &h) {
  const R *a = rec + t;
  h.r = a[R3_R * vs];     h.pp = a[R3_P * vs];    h.u = a[R3_U * vs];     h.v = a[R3_V * vs];
  h.w = a[R3_W * vs];     h.A = a[R3_A * vs];     h.B = a[R3_B * vs];     h.C = a[R3_C * vs];
  h.drx = a[R3_DRX * vs]; h.dpx = a[R3_DPX * vs]; h.dux = a[R3_DUX * vs]; h.dvx = a[R3_DVX * vs];
  h.dwx = a[R3_DWX * vs]; h.dBx = a[R3_DBX * vs]; h.dCx = a[R3_DCX * vs];
  h.dry = a[R3_DRY * vs]; h.dpy = a[R3_DPY * vs]; h.duy = a[R3_DUY * vs]; h.dvy = a[R3_DVY * vs];
  h.dwy = a[R3_DWY * vs]; h.dAy = a[R3_DAY * vs]; h.dCy = a[R3_DCY * vs];
  h.drz = a[R3_DRZ * vs]; h.dpz = a[R3_DPZ * vs]; h.duz = a[R3_DUZ * vs]; h.dvz = a[R3_DVZ * vs];
  h.dwz = a[R3_DWZ * vs]; h.dAz = a[R3_DAZ * vs]; h.dBz = a[R3_DBZ * vs];
  h.AL = a[R3_AL * vs];   h.BL = a[R3_BL * vs];   h.CL = a[R3_CL * vs];
  h.dALy = a[R3_DALY * vs]; h.dALz = a[R3_DALZ * vs]; h.dBLx = a[R3_DBLX * vs];
  h.dBLz = a[R3_DBLZ * vs]; h.dCLx = a[R3_DCLX * vs]; h.dCLy = a[R3_DCLY * vs];
  h.AR = h.BR = h.CR = T(0);
  h.dARy = h.dARz = h.dBRx = h.dBRz = h.dCRx = h.dCRy = T(0);
}

template <class R>
PPK_HD void store_rec2(R *__restrict__ rec, long long vs, long long t, const Half2D<R> &h) {
  R *o = rec + t;
  o[R2_R * vs] = h.r;     o[R2_P * vs] = h.pp;    o[R2_U * vs] = h.u;     o[R2_V * vs] = h.v;
  o[R2_W * vs] = h.w;     o[R2_A * vs] = h.A;     o[R2_B * vs] = h.B;     o[R2_C * vs] = h.C;
  o[R2_DRX * vs] = h.drx; o[R2_DPX * vs] = h.dpx; o[R2_DUX * vs] = h.dux; o[R2_DVX * vs] = h.dvx;
  o[R2_DWX * vs] = h.dwx; o[R2_DBX * vs] = h.dBx; o[R2_DCX * vs] = h.dCx;
  o[R2_DRY * vs] = h.dry; o[R2_DPY * vs] = h.dpy; o[R2_DUY * vs] = h.duy; o[R2_DVY * vs] = h.dvy;
  o[R2_DWY * vs] = h.dwy; o[R2_DAY * vs] = h.dAy; o[R2_DCY * vs] = h.dCy;
  o[R2_AL * vs] = h.AL;   o[R2_BL * vs] = h.BL;
  o[R2_DALY * vs] = h.dALy; o[R2_DBLX * vs] = h.dBLx;
}

template <class R, class T>
PPK_HD void load_rec2(const R *__restrict__ rec, long long vs, long long t, Half2D<T> &h) {
  const R *a = rec + t;
  h.r = a[R2_R * vs];     h.pp = a[R2_P * vs];    h.u = a[R2_U * vs];     h.v = a[R2_V * vs];
  h.w = a[R2_W * vs];     h.A = a[R2_A * vs];     h.B = a[R2_B * vs];     h.C = a[R2_C * vs];
  h.drx = a[R2_DRX * vs]; h.dpx = a[R2_DPX * vs]; h.dux = a[R2_DUX * vs]; h.dvx = a[R2_DVX * vs];
  h.dwx = a[R2_DWX * vs]; h.dBx = a[R2_DBX * vs]; h.dCx = a[R2_DCX * vs];
  h.dry = a[R2_DRY * vs]; h.dpy = a[R2_DPY * vs]; h.duy = a[R2_DUY * vs]; h.dvy = a[R2_DVY * vs];
  h.dwy = a[R2_DWY * vs]; h.dAy = a[R2_DAY * vs]; h.dCy = a[R2_DCY * vs];
  h.AL = a[R2_AL * vs];   h.BL = a[R2_BL * vs];
  h.dALy = a[R2_DALY * vs]; h.dBLx = a[R2_DBLX * vs];
  h.AR = h.BR = T(0);
  h.dARy = h.dBRx = T(0);
}

// ===========================================================================================
// half step, 3D.  A tile covers TX x TY x TZ record cells starting at b0 (record cells span
// [gw-1, size-gw] per direction).  LDS holds, with origin b0-1, the face fields over
// (T+3)^3 and the primitives over (T+2)^3, and with origin b0 the edge E field over (T+1)^3;
// entries beyond the last cell any record needs are neither loaded nor read.
// ===========================================================================================
template <int TX_, int TY_, int TZ_>
struct HalfTile3 {
  static constexpr int TX = TX_, TY = TY_, TZ = TZ_, NT = TX * TY * TZ;
  static constexpr int FX = TX + 3, FY = TY + 3, FZ = TZ + 3, NF = FX * FY * FZ;
  static constexpr int QX = TX + 2, QY = TY + 2, QZ = TZ + 2, NQ = QX * QY * QZ;
  static constexpr int EX = TX + 1, EY = TY + 1, EZ = TZ + 1, NE = EX * EY * EZ;
  static constexpr int LDS = 3 * NF + 8 * NQ + 3 * NE;  // elements of the arithmetic type
  PPK_HD static int f(int x, int y, int z) { return x + FX * (y + FY * z); }
  PPK_HD static int q(int x, int y, int z) { return x + QX * (y + QY * z); }
  PPK_HD static int e(int x, int y, int z) { return x + EX * (y + EY * z); }
};
// compiled 3D tile shapes: Tile3 for single grids, all three for the batch's autotuner
using Tile3 = HalfTile3<16, 4, 4>;   // 256 threads, 74 KB LDS
using Tile3d = HalfTile3<8, 8, 4>;   // 256 threads, 68 KB
using Tile3e = HalfTile3<32, 4, 2>;  // 256 threads, 94 KB
static_assert(Tile3e::LDS * sizeof(double) <= 160 * 1024, "3D tile exceeds the LDS of a CU");

// the global loads of phase 1, held in registers until stashed in LDS (so a persistent block
// can issue the next tile's loads before it computes the current tile's records)
template <class T>
struct HalfFetch3 {
  static constexpr int IF = (T::NF + T::NT - 1) / T::NT, IQ = (T::NQ + T::NT - 1) / T::NT;
  double fv[IF][3], qv[IQ][5];
  bool fok[IF], qok[IQ];
};
template <class T>
PPK_HD void half3_fetch(const double *__restrict__ U, const DevParams &p, int b0x, int b0y,
                        int b0z, int tid, HalfFetch3<T> &f) {
  constexpr int IF = HalfFetch3<T>::IF, IQ = HalfFetch3<T>::IQ;
  const long long vs = p.vstride;
  const int hi = 1 - p.gw;  // last primitive cell: size + hi
#pragma unroll
  for (int it = 0; it < IF; ++it) {
    const int n = tid + it * T::NT;
    const int x = n % T::FX, r = n / T::FX, y = r % T::FY, z = r / T::FY;
    const int i = b0x - 1 + x, j = b0y - 1 + y, k = b0z - 1 + z;
    f.fok[it] = n < T::NF && i < p.isize && j < p.jsize && k < p.ksize;
    // unconditional load at a clamped (in-box) address keeps all loads in one batch
    const long long t = cidx(p, min(i, p.isize - 1), min(j, p.jsize - 1), min(k, p.ksize - 1));
    f.fv[it][0] = U[t + IA * vs];
    f.fv[it][1] = U[t + IB * vs];
    f.fv[it][2] = U[t + IC * vs];
  }
#pragma unroll
  for (int it = 0; it < IQ; ++it) {
    const int n = tid + it * T::NT;
    const int x = n % T::QX, r = n / T::QX, y = r % T::QY, z = r / T::QY;
    const int i = b0x - 1 + x, j = b0y - 1 + y, k = b0z - 1 + z;
    f.qok[it] = n < T::NQ && i <= p.isize + hi && j <= p.jsize + hi && k <= p.ksize + hi;
    const long long t =
        cidx(p, min(i, p.isize + hi), min(j, p.jsize + hi), min(k, p.ksize + hi));
#pragma unroll
    for (int v = ID; v <= IW; ++v) f.qv[it][v] = U[t + v * vs];
  }
}
template <class T, class R>
PPK_HD void half3_stash(const HalfFetch3<T> &f, int tid, R *sF, R *sQ) {
  constexpr int IF = HalfFetch3<T>::IF, IQ = HalfFetch3<T>::IQ;
#pragma unroll
  for (int it = 0; it < IF; ++it) {
    const int n = tid + it * T::NT;
    if (f.fok[it]) {
      sF[n] = R(f.fv[it][0]);
      sF[T::NF + n] = R(f.fv[it][1]);
      sF[2 * T::NF + n] = R(f.fv[it][2]);
    }
  }
#pragma unroll
  for (int it = 0; it < IQ; ++it) {
    const int n = tid + it * T::NT;
    if (f.qok[it]) {
#pragma unroll
      for (int v = ID; v <= IW; ++v) sQ[v * T::NQ + n] = R(f.qv[it][v]);
    }
  }
}
// phase 1, all global loads of the tile: face fields A, B, C of cells [b0-1, b0+T+1]
// (clipped to the box) into sF, and the five other conservative variables of the primitive
// cells [b0-1, b0+T] (up to size-gw+1) into their slots of sQ.  Every load of a thread is
// issued before its first LDS store, so the tile pays one memory latency, not one per
// strided pass.
template <class T, class R>
PPK_HD void half3_load(const double *__restrict__ U, const DevParams &p, int b0x, int b0y,
                       int b0z, int tid, R *sF, R *sQ) {
  HalfFetch3<T> f;
  half3_fetch<T>(U, p, b0x, b0y, b0z, tid, f);
  half3_stash<T, R>(f, tid, sF, sQ);
}

// phase 2: primitives of cells [b0-1, b0+T] up to size-gw+1 (MHDRunFunctors3D.h:88-163), in
// place: each thread reads the raw slots of its own cells only
template <class T, class R>
PPK_HD void half3_prims(const DevParams &p, int b0x, int b0y, int b0z, int tid, const R *sF,
                        R *sQ) {
  const int hi = 1 - p.gw;
  for (int n = tid; n < T::NQ; n += T::NT) {
    const int x = n % T::QX, r = n / T::QX, y = r % T::QY, z = r / T::QY;
    const int i = b0x - 1 + x, j = b0y - 1 + y, k = b0z - 1 + z;
    if (i > p.isize + hi || j > p.jsize + hi || k > p.ksize + hi) continue;
    St<R> u, q;
#pragma unroll
    for (int v = ID; v <= IW; ++v) u[v] = sQ[v * T::NQ + n];
    const int f0 = T::f(x, y, z);
    u[IA] = sF[f0];
    u[IB] = sF[T::NF + f0];
    u[IC] = sF[2 * T::NF + f0];
    const R mx = sF[T::f(x + 1, y, z)];
    const R my = sF[T::NF + T::f(x, y + 1, z)];
    const R mz = sF[2 * T::NF + T::f(x, y, z + 1)];
    constoprim(u, mx, my, mz, q, p);
#pragma unroll
    for (int v = 0; v < 8; ++v) sQ[v * T::NQ + n] = q[v];
  }
}

// phase 3: edge electric field E = v x B of cells [b0, b0+T] up to size-gw+1, same
// operations as the staged E kernel (MHDRunFunctors3D.h:303-356)
template <class T, class R>
PPK_HD void half3_elec(const DevParams &p, int b0x, int b0y, int b0z, int tid, const R *sF,
                       const R *sQ, R *sE) {
  const int hi = 1 - p.gw;
  for (int n = tid; n < T::NE; n += T::NT) {
    const int x = n % T::EX, r = n / T::EX, y = r % T::EY, z = r / T::EY;
    if (b0x + x > p.isize + hi || b0y + y > p.jsize + hi || b0z + z > p.ksize + hi) continue;
    const int X = x + 1, Y = y + 1, Z = z + 1;  // position in the face / primitive tiles
    auto q = [&](int dx, int dy, int dz, int v) {
      return sQ[v * T::NQ + T::q(X + dx, Y + dy, Z + dz)];
    };
    auto u = [&](int dx, int dy, int dz, int c) {
      return sF[c * T::NF + T::f(X + dx, Y + dy, Z + dz)];
    };
    R v_, w_, u_, A, B, C;
    v_ = R(0.25) * (q(0, -1, -1, IV) + q(0, -1, 0, IV) + q(0, 0, -1, IV) + q(0, 0, 0, IV));
    w_ = R(0.25) * (q(0, -1, -1, IW) + q(0, -1, 0, IW) + q(0, 0, -1, IW) + q(0, 0, 0, IW));
    B = R(0.5) * (u(0, 0, -1, 1) + u(0, 0, 0, 1));
    C = R(0.5) * (u(0, -1, 0, 2) + u(0, 0, 0, 2));
    sE[n] = v_ * C - w_ * B;
    u_ = R(0.25) * (q(-1, 0, -1, IU) + q(-1, 0, 0, IU) + q(0, 0, -1, IU) + q(0, 0, 0, IU));
    w_ = R(0.25) * (q(-1, 0, -1, IW) + q(-1, 0, 0, IW) + q(0, 0, -1, IW) + q(0, 0, 0, IW));
    A = R(0.5) * (u(0, 0, -1, 0) + u(0, 0, 0, 0));
    C = R(0.5) * (u(-1, 0, 0, 2) + u(0, 0, 0, 2));
    sE[T::NE + n] = w_ * A - u_ * C;
    u_ = R(0.25) * (q(-1, -1, 0, IU) + q(-1, 0, 0, IU) + q(0, -1, 0, IU) + q(0, 0, 0, IU));
    v_ = R(0.25) * (q(-1, -1, 0, IV) + q(-1, 0, 0, IV) + q(0, -1, 0, IV) + q(0, 0, 0, IV));
    A = R(0.5) * (u(0, -1, 0, 0) + u(0, 0, 0, 0));
    B = R(0.5) * (u(-1, 0, 0, 1) + u(0, 0, 0, 1));
    sE[2 * T::NE + n] = u_ * B - v_ * A;
  }
}

// phase 4: one record cell per thread: slopes, transverse face-field slopes
// (MHDRunFunctors3D.h:472-532) and the half step (MHDBaseFunctor3D.h:688-1114)
template <class T, class R>
PPK_HD void half3_record(const DevParams &p, int b0x, int b0y, int b0z, int tid, const R *sF,
                         const R *sQ, const R *sE, double dt, R *__restrict__ rec) {
  const int x = tid % T::TX, y = (tid / T::TX) % T::TY, z = tid / (T::TX * T::TY);
  const int i = b0x + x, j = b0y + y, k = b0z + z;
  if (i > p.isize - p.gw || j > p.jsize - p.gw || k > p.ksize - p.gw) return;
  const int X = x + 1, Y = y + 1, Z = z + 1;
  auto ld = [&](int dx, int dy, int dz, St<R> &s) {
    const int n = T::q(X + dx, Y + dy, Z + dz);
#pragma unroll
    for (int v = 0; v < 8; ++v) s[v] = sQ[v * T::NQ + n];
  };
  St<R> q, qPX, qMX, qPY, qMY, qPZ, qMZ;
  ld(0, 0, 0, q);
  ld(1, 0, 0, qPX);
  ld(-1, 0, 0, qMX);
  ld(0, 1, 0, qPY);
  ld(0, -1, 0, qMY);
  ld(0, 0, 1, qPZ);
  ld(0, 0, -1, qMZ);
  auto F = [&](int dx, int dy, int dz, int c) {
    return sF[c * T::NF + T::f(X + dx, Y + dy, Z + dz)];
  };
  const R bfNb[6] = {F(0, 0, 0, 0), F(1, 0, 0, 0), F(0, 0, 0, 1),
                          F(0, 1, 0, 1), F(0, 0, 0, 2), F(0, 0, 1, 2)};
  const R xst = R(fmin(p.slope_type, 2.0));
  // the left-side transverse slopes; the right-side ones (second half) are never used
  const R d12[12] = {limited_slope(xst, F(0, 0, 0, 0), F(0, 1, 0, 0), F(0, -1, 0, 0)),
                          limited_slope(xst, F(0, 0, 0, 0), F(0, 0, 1, 0), F(0, 0, -1, 0)),
                          limited_slope(xst, F(0, 0, 0, 1), F(1, 0, 0, 1), F(-1, 0, 0, 1)),
                          limited_slope(xst, F(0, 0, 0, 1), F(0, 0, 1, 1), F(0, 0, -1, 1)),
                          limited_slope(xst, F(0, 0, 0, 2), F(1, 0, 0, 2), F(-1, 0, 0, 2)),
                          limited_slope(xst, F(0, 0, 0, 2), F(0, 1, 0, 2), F(0, -1, 0, 2)),
                          R(0), R(0), R(0), R(0), R(0), R(0)};
  auto E = [&](int dx, int dy, int dz, int c) {
    return sE[c * T::NE + T::e(x + dx, y + dy, z + dz)];
  };
  // only the fields of the left faces' edges enter the record (ERR, FRR, GRR do not)
  const R Ex[4] = {E(0, 0, 0, 0), E(0, 0, 1, 0), E(0, 1, 0, 0), R(0)};
  const R Ey[4] = {E(0, 0, 0, 1), E(0, 0, 1, 1), E(1, 0, 0, 1), R(0)};
  const R Ez[4] = {E(0, 0, 0, 2), E(0, 1, 0, 2), E(1, 0, 0, 2), R(0)};
  const R dtdx = R(dt / p.dx), dtdy = R(dt / p.dy), dtdz = R(dt / p.dz);
  Half3D<R> h;
  half_step_3d(q, qPX, qMX, qPY, qMY, qPZ, qMZ, bfNb, d12, Ex, Ey, Ez, dtdx, dtdy, dtdz, p, h);
  store_rec3(rec, p.vstride, cidx(p, i, j, k), h);
}

// ===========================================================================================
// half step, 2D (same scheme; face fields A, B; one E component at cell corners)
// ===========================================================================================
template <int TX_, int TY_>
struct HalfTile2 {
  static constexpr int TX = TX_, TY = TY_, NT = TX * TY;
  static constexpr int FX = TX + 3, FY = TY + 3, NF = FX * FY;
  static constexpr int QX = TX + 2, QY = TY + 2, NQ = QX * QY;
  static constexpr int EX = TX + 1, EY = TY + 1, NE = EX * EY;
  static constexpr int LDS = 2 * NF + 8 * NQ + NE;
  PPK_HD static int f(int x, int y) { return x + FX * y; }
  PPK_HD static int q(int x, int y) { return x + QX * y; }
  PPK_HD static int e(int x, int y) { return x + EX * y; }
};
using Tile2 = HalfTile2<32, 8>;   // 256 threads, 30 KB LDS
using Tile2b = HalfTile2<64, 4>;  // 256 threads, 35 KB
using Tile2c = HalfTile2<16, 16>; // 256 threads, 29 KB

// all global loads of the tile: face fields A, B into sF, the six cell-centred variables
// (ID..IW and the 2D IC) of the primitive cells into their slots of sQ; loads first, then
// the LDS stores
template <class T, class R>
PPK_HD void half2_load(const double *__restrict__ U, const DevParams &p, int b0x, int b0y, int tid,
                       R *sF, R *sQ) {
  constexpr int IF = (T::NF + T::NT - 1) / T::NT, IQ = (T::NQ + T::NT - 1) / T::NT;
  const long long vs = p.vstride;
  const int hi = 1 - p.gw;
  double fv[IF][2], qv[IQ][6];
  bool fok[IF], qok[IQ];
#pragma unroll
  for (int it = 0; it < IF; ++it) {
    const int n = tid + it * T::NT;
    const int x = n % T::FX, y = n / T::FX;
    const int i = b0x - 1 + x, j = b0y - 1 + y;
    fok[it] = n < T::NF && i < p.isize && j < p.jsize;
    const long long t = cidx(p, min(i, p.isize - 1), min(j, p.jsize - 1), 0);
    fv[it][0] = U[t + IA * vs];
    fv[it][1] = U[t + IB * vs];
  }
#pragma unroll
  for (int it = 0; it < IQ; ++it) {
    const int n = tid + it * T::NT;
    const int x = n % T::QX, y = n / T::QX;
    const int i = b0x - 1 + x, j = b0y - 1 + y;
    qok[it] = n < T::NQ && i <= p.isize + hi && j <= p.jsize + hi;
    const long long t = cidx(p, min(i, p.isize + hi), min(j, p.jsize + hi), 0);
#pragma unroll
    for (int v = ID; v <= IW; ++v) qv[it][v] = U[t + v * vs];
    qv[it][5] = U[t + IC * vs];
  }
#pragma unroll
  for (int it = 0; it < IF; ++it) {
    const int n = tid + it * T::NT;
    if (fok[it]) {
      sF[n] = R(fv[it][0]);
      sF[T::NF + n] = R(fv[it][1]);
    }
  }
#pragma unroll
  for (int it = 0; it < IQ; ++it) {
    const int n = tid + it * T::NT;
    if (qok[it]) {
#pragma unroll
      for (int v = ID; v <= IW; ++v) sQ[v * T::NQ + n] = R(qv[it][v]);
      sQ[IC * T::NQ + n] = R(qv[it][5]);
    }
  }
}

// primitives in place (MHDRunFunctors2D.h ComputePrimitivesFunctor): Bz is cell-centred in 2D
// and is averaged with 0
template <class T, class R>
PPK_HD void half2_prims(const DevParams &p, int b0x, int b0y, int tid, const R *sF, R *sQ) {
  const int hi = 1 - p.gw;
  for (int n = tid; n < T::NQ; n += T::NT) {
    const int x = n % T::QX, y = n / T::QX;
    const int i = b0x - 1 + x, j = b0y - 1 + y;
    if (i > p.isize + hi || j > p.jsize + hi) continue;
    St<R> u, q;
#pragma unroll
    for (int v = ID; v <= IW; ++v) u[v] = sQ[v * T::NQ + n];
    u[IC] = sQ[IC * T::NQ + n];
    const int f0 = T::f(x, y);
    u[IA] = sF[f0];
    u[IB] = sF[T::NF + f0];
    const R mx = sF[T::f(x + 1, y)];
    const R my = sF[T::NF + T::f(x, y + 1)];
    constoprim(u, mx, my, R(0), q, p);
#pragma unroll
    for (int v = 0; v < 8; ++v) sQ[v * T::NQ + n] = q[v];
  }
}

// E_z at the lower-left corner of cells [b0, b0+T] up to size-gw+1
template <class T, class R>
PPK_HD void half2_elec(const DevParams &p, int b0x, int b0y, int tid, const R *sF, const R *sQ,
                       R *sE) {
  const int hi = 1 - p.gw;
  for (int n = tid; n < T::NE; n += T::NT) {
    const int x = n % T::EX, y = n / T::EX;
    if (b0x + x > p.isize + hi || b0y + y > p.jsize + hi) continue;
    const int X = x + 1, Y = y + 1;
    auto q = [&](int dx, int dy, int v) { return sQ[v * T::NQ + T::q(X + dx, Y + dy)]; };
    auto u = [&](int dx, int dy, int c) { return sF[c * T::NF + T::f(X + dx, Y + dy)]; };
    sE[n] = elec_z_2d(q(-1, -1, IU), q(-1, 0, IU), q(0, -1, IU), q(0, 0, IU), q(-1, -1, IV),
                      q(-1, 0, IV), q(0, -1, IV), q(0, 0, IV), u(0, -1, 0), u(0, 0, 0),
                      u(-1, 0, 1), u(0, 0, 1));
  }
}

// one record cell per thread (MHDBaseFunctor2D.h:774-1109)
template <class T, class R>
PPK_HD void half2_record(const DevParams &p, int b0x, int b0y, int tid, const R *sF, const R *sQ,
                         const R *sE, double dt, R *__restrict__ rec) {
  const int x = tid % T::TX, y = tid / T::TX;
  const int i = b0x + x, j = b0y + y;
  if (i > p.isize - p.gw || j > p.jsize - p.gw) return;
  St<R> qNb[3][3];
  R bf[4][4][3];
#pragma unroll
  for (int di = 0; di < 3; ++di)
#pragma unroll
    for (int dj = 0; dj < 3; ++dj) {
      const int n = T::q(x + di, y + dj);
#pragma unroll
      for (int v = 0; v < 8; ++v) qNb[di][dj][v] = sQ[v * T::NQ + n];
    }
#pragma unroll
  for (int di = 0; di < 4; ++di)
#pragma unroll
    for (int dj = 0; dj < 4; ++dj) {
      const int n = T::f(x + di, y + dj);
      bf[di][dj][0] = sF[n];
      bf[di][dj][1] = sF[T::NF + n];
      bf[di][dj][2] = R(0);
    }
  const R ELL = sE[T::e(x, y)], ELR = sE[T::e(x, y + 1)];
  const R ERL = sE[T::e(x + 1, y)], ERR = sE[T::e(x + 1, y + 1)];
  Half2D<R> h;
  half_step_2d(qNb, bf, ELL, ELR, ERL, ERR, R(dt / p.dx), R(dt / p.dy), p, h);
  store_rec2(rec, p.vstride, cidx(p, i, j, 0), h);
}

// ===========================================================================================
// Riemann problems at the lower faces and EMFs at the lower edges of cell (i,j,k), for cells
// [gw, size-gw] per direction (the ranges of the reference's flux and EMF functors)
// ===========================================================================================
template <class R>
PPK_HD void swap_dir(St<R> &s, int a, int b) {
  R t = s[a];
  s[a] = s[b];
  s[b] = t;
}

// One face of cell c = (i,j,k) (its lower face in direction D): qm of cell c-D (whose right
// face field is the left one of c) against qp of cell c, rotated so that the normal is x,
// then the 1D Riemann solver; the 5 (3D) or 6 (2D) flux components are stored rotated.
// Reference src/muscl/MHDRunFunctors3D.h:1839-1902 and MHDRunFunctors2D.h.
template <int D, class R, class T>
PPK_HD void face_states3(const R *__restrict__ rec, const DevParams &p, long long c, St<T> &ql,
                         St<T> &qr) {
  const long long vs = p.vstride;
  const long long sd = D == 0 ? 1 : (D == 1 ? (long long)p.isize : p.plane);
  Half3D<T> hl, hr;
  load_rec3(rec, vs, c - sd, hl);
  load_rec3(rec, vs, c, hr);
  if (D == 0) hl.AR = rec[c + R3_AL * vs];
  if (D == 1) hl.BR = rec[c + R3_BL * vs];
  if (D == 2) hl.CR = rec[c + R3_CL * vs];
  ql = face_state_3d(hl, D, +1, p);
  qr = face_state_3d(hr, D, -1, p);
  if (D == 1) {
    swap_dir(ql, IU, IV); swap_dir(ql, IA, IB);
    swap_dir(qr, IU, IV); swap_dir(qr, IA, IB);
  } else if (D == 2) {
    swap_dir(ql, IU, IW); swap_dir(ql, IA, IC);
    swap_dir(qr, IU, IW); swap_dir(qr, IA, IC);
  }
}
template <int D, class R>
PPK_HD void store_flux3(R *__restrict__ flux, const DevParams &p, long long c, const St<R> &f) {
  const long long vs = p.vstride;
  R *o = flux + (long long)D * kFluxComp3 * vs + c;
  o[0] = f[ID];
  o[vs] = f[IP];
  o[2 * vs] = f[IU];
  o[3 * vs] = f[IV];
  o[4 * vs] = f[IW];
}
template <int D, int RS = -1, class R>
PPK_HD void face3(const R *__restrict__ rec, R *__restrict__ flux, const DevParams &p,
                  long long c) {
  St<R> ql, qr, f;
  face_states3<D>(rec, p, c, ql, qr);
  riemann_mhd_rs<RS>(ql, qr, f, p);
  store_flux3<D>(flux, p, c, f);
}

// One edge of cell c (its lower edge along direction E = 0 x, 1 y, 2 z): the four edge states
// of the cells around it, then compute_emf.  Reference src/muscl/MHDRunFunctors3D.h:2183-2230
// (the y edge takes RB and LT swapped, :2212-2217).  The edge states and the EMF solve are
// double whatever the record type (the FP32-bulk form keeps the CT EMFs in FP64).
template <int E, class R>
PPK_HD void edge3(const R *__restrict__ rec, double *__restrict__ emf, const DevParams &p,
                  long long c) {
  const long long vs = p.vstride, sx = 1, sy = p.isize, sz = p.plane;
  auto RF = [&](long long t, int f) { return double(rec[t + f * vs]); };
  Half3D<double> a, b, d, h;
  load_rec3(rec, vs, c, h);
  if (E == 2) {  // RT of c-x-y, RB of c-x, LT of c-y, LB of c
    load_rec3(rec, vs, c - sx - sy, a);
    a.AR = RF(c - sy, R3_AL); a.dARy = RF(c - sy, R3_DALY);
    a.BR = RF(c - sx, R3_BL); a.dBRx = RF(c - sx, R3_DBLX);
    load_rec3(rec, vs, c - sx, b);
    b.AR = h.AL; b.dARy = h.dALy;
    load_rec3(rec, vs, c - sy, d);
    d.BR = h.BL; d.dBRx = h.dBLx;
    emf[c + I_EMFZ * vs] =
        compute_emf<EMFZ>(edge_state_3d(a, 2, IRT, p), edge_state_3d(b, 2, IRB, p),
                          edge_state_3d(d, 2, ILT, p), edge_state_3d(h, 2, ILB, p), p);
  } else if (E == 1) {  // RT of c-x-z, LT of c-z, RB of c-x, LB of c
    load_rec3(rec, vs, c - sx - sz, a);
    a.AR = RF(c - sz, R3_AL); a.dARz = RF(c - sz, R3_DALZ);
    a.CR = RF(c - sx, R3_CL); a.dCRx = RF(c - sx, R3_DCLX);
    load_rec3(rec, vs, c - sz, b);
    b.CR = h.CL; b.dCRx = h.dCLx;
    load_rec3(rec, vs, c - sx, d);
    d.AR = h.AL; d.dARz = h.dALz;
    emf[c + I_EMFY * vs] =
        compute_emf<EMFY>(edge_state_3d(a, 1, IRT, p), edge_state_3d(b, 1, ILT, p),
                          edge_state_3d(d, 1, IRB, p), edge_state_3d(h, 1, ILB, p), p);
  } else {  // RT of c-y-z, RB of c-y, LT of c-z, LB of c
    load_rec3(rec, vs, c - sy - sz, a);
    a.BR = RF(c - sz, R3_BL); a.dBRz = RF(c - sz, R3_DBLZ);
    a.CR = RF(c - sy, R3_CL); a.dCRy = RF(c - sy, R3_DCLY);
    load_rec3(rec, vs, c - sy, b);
    b.BR = h.BL; b.dBRz = h.dBLz;
    load_rec3(rec, vs, c - sz, d);
    d.CR = h.CL; d.dCRy = h.dCLy;
    emf[c + I_EMFX * vs] =
        compute_emf<EMFX>(edge_state_3d(a, 0, IRT, p), edge_state_3d(b, 0, IRB, p),
                          edge_state_3d(d, 0, ILT, p), edge_state_3d(h, 0, ILB, p), p);
  }
}

// 2D faces (D = 0, 1) and the z edge, reference src/muscl/MHDRunFunctors2D.h
// ComputeFluxesAndStoreFunctor and ComputeEmfAndStoreFunctor (:654-680)
template <int D, class R, class T>
PPK_HD void face_states2(const R *__restrict__ rec, const DevParams &p, long long c, St<T> &ql,
                         St<T> &qr) {
  const long long vs = p.vstride;
  const long long sd = D == 0 ? 1 : (long long)p.isize;
  Half2D<T> hl, hr;
  load_rec2(rec, vs, c - sd, hl);
  load_rec2(rec, vs, c, hr);
  if (D == 0) hl.AR = rec[c + R2_AL * vs];
  if (D == 1) hl.BR = rec[c + R2_BL * vs];
  ql = face_state_2d(hl, D, +1, p);
  qr = face_state_2d(hr, D, -1, p);
  if (D == 1) {
    swap_dir(ql, IU, IV); swap_dir(ql, IA, IB);
    swap_dir(qr, IU, IV); swap_dir(qr, IA, IB);
  }
}
template <int D, class R>
PPK_HD void store_flux2(R *__restrict__ flux, const DevParams &p, long long c, const St<R> &f) {
  const long long vs = p.vstride;
  R *o = flux + (long long)D * kFluxComp2 * vs + c;
  o[0] = f[ID];
  o[vs] = f[IP];
  o[2 * vs] = f[IU];
  o[3 * vs] = f[IV];
  o[4 * vs] = f[IW];
  o[5 * vs] = f[IC];
}
template <int D, int RS = -1, class R>
PPK_HD void face2(const R *__restrict__ rec, R *__restrict__ flux, const DevParams &p,
                  long long c) {
  St<R> ql, qr, f;
  face_states2<D>(rec, p, c, ql, qr);
  riemann_mhd_rs<RS>(ql, qr, f, p);
  store_flux2<D>(flux, p, c, f);
}

// (double edge states and EMF whatever the record type, as edge3)
template <class R>
PPK_HD void edge2(const R *__restrict__ rec, double *__restrict__ emf, const DevParams &p,
                  long long c) {
  const long long vs = p.vstride, sx = 1, sy = p.isize;
  auto RF = [&](long long t, int f) { return double(rec[t + f * vs]); };
  Half2D<double> a, b, d, h;
  load_rec2(rec, vs, c, h);
  load_rec2(rec, vs, c - sx - sy, a);
  a.AR = RF(c - sy, R2_AL); a.dARy = RF(c - sy, R2_DALY);
  a.BR = RF(c - sx, R2_BL); a.dBRx = RF(c - sx, R2_DBLX);
  load_rec2(rec, vs, c - sx, b);
  b.AR = h.AL; b.dARy = h.dALy;
  load_rec2(rec, vs, c - sy, d);
  d.BR = h.BL; d.dBRx = h.dBLx;
  emf[c] = compute_emf<EMFZ>(edge_state_2d(a, IRT, p), edge_state_2d(b, IRB, p),
                             edge_state_2d(d, ILT, p), edge_state_2d(h, ILB, p), p);
}

// cell (i,j,k) of thread tid of brick L; false past the covered range
// row r of an enumeration of (j, k) rows in patches of pj x pk rows, npj patches along j.
// Cell and row numbers are below 2^32 (grids are limited to 2^31 padded cells), so the
// decoding is 32-bit unsigned arithmetic: a 64-bit division by a run-time value costs
// several times more instructions, and every thread of the per-cell kernels decodes its cell.
PPK_HD void patch_row(unsigned r, int pj, int pk, int npj, int &j, int &k) {
  unsigned lj, lk;
  if (pj == kPatch && pk == kPatch) {  // the 3D patch: shifts
    static_assert(kPatch == 4, "patch decoding by shifts assumes 4 x 4 row patches");
    lj = r & 3u;
    lk = (r >> 2) & 3u;
    r >>= 4;
  } else {
    lj = r % (unsigned)pj;
    r /= (unsigned)pj;
    lk = r % (unsigned)pk;
    r /= (unsigned)pk;
  }
  const unsigned pr = r / (unsigned)npj;
  j = (int)(r - pr * (unsigned)npj) * pj + (int)lj;
  k = (int)pr * pk + (int)lk;
}

PPK_HD bool brick_cell(const Brick &b, int L, int tid, int &i, int &j, int &k) {
  if (b.bx == 0) {  // flattened runs of 256 cells, rows in patches
    const unsigned n = (unsigned)L * 256u + (unsigned)tid;
    const unsigned row = n / (unsigned)b.mx;
    i = (int)(n - row * (unsigned)b.mx);
    patch_row(row, b.by, b.bz, b.ny, j, k);
    return j < b.my && k < b.mz && row < (unsigned)(b.ny * b.by) * (unsigned)(b.nz * b.bz);
  }
  const int ux = L % b.nx, r = L / b.nx, uy = r % b.ny, uz = r / b.ny;
  i = ux * b.bx + tid % b.bx;
  j = uy * b.by + (tid / b.bx) % b.by;
  k = uz * b.bz + tid / (b.bx * b.by);
  return i < b.mx && j < b.my && k < b.mz;
}

// padded-box index of thread tid of sub-box brick L, or -1; kk: its z index in the sub-box
PPK_HD long long sub_brick_cell(const DevParams &p, const Brick &b, int L, int tid, int &kk) {
  int i, j, k;
  if (!brick_cell(b, L, tid, i, j, k)) return -1;
  kk = p.dim == 3 ? k : 0;
  return cidx(p, p.gw + i, p.gw + j, p.dim == 3 ? p.gw + k : 0);
}
PPK_HD long long sub_brick_cell(const DevParams &p, const Brick &b, int L, int tid) {
  int kk;
  return sub_brick_cell(p, b, L, tid, kk);
}

// ===========================================================================================
// update + CT + ghost copy + CFL of the new state
// ===========================================================================================
// CT of one face field (MHDRunFunctors3D.h:2600-2620); each component in the reference's order
PPK_HD double ct3_a(double a, const double *__restrict__ emf, long long vs, long long t,
                    long long sy, long long sz, double dtdy, double dtdz) {
  a += (double(emf[t + sy + I_EMFZ * vs]) - double(emf[t + I_EMFZ * vs])) * dtdy;
  a -= (double(emf[t + sz + I_EMFY * vs]) - double(emf[t + I_EMFY * vs])) * dtdz;
  return a;
}
PPK_HD double ct3_b(double b, const double *__restrict__ emf, long long vs, long long t,
                    long long sx, long long sz, double dtdx, double dtdz) {
  b -= (double(emf[t + sx + I_EMFZ * vs]) - double(emf[t + I_EMFZ * vs])) * dtdx;
  b += (double(emf[t + sz + I_EMFX * vs]) - double(emf[t + I_EMFX * vs])) * dtdz;
  return b;
}
PPK_HD double ct3_c(double c, const double *__restrict__ emf, long long vs, long long t,
                    long long sx, long long sy, double dtdx, double dtdy) {
  c += (double(emf[t + sx + I_EMFY * vs]) - double(emf[t + I_EMFY * vs])) * dtdx;
  c -= (double(emf[t + sy + I_EMFX * vs]) - double(emf[t + I_EMFX * vs])) * dtdy;
  return c;
}
// 2D CT (MHDRunFunctors2D.h:1690-1691)
PPK_HD double ct2_a(double a, const double *__restrict__ emf, long long t, long long sy,
                    double dtdy) {
  a += (double(emf[t + sy]) - double(emf[t])) * dtdy;
  return a;
}
PPK_HD double ct2_b(double b, const double *__restrict__ emf, long long t, long long sx,
                    double dtdx) {
  b -= (double(emf[t + sx]) - double(emf[t])) * dtdx;
  return b;
}

// new state of padded-box cell t (ghosts copied), written to Uout; returns the cell's CFL
// term sum_d (c_f + |v_d|)/dx_d when fold is set and the cell is interior, else 0.  The
// primitive of an interior cell needs the new face field of its +1 neighbour: the CT value of
// that face, or past the last interior cell the value the next ghost fill puts there
// (periodic image, or +-own field, BoundariesFunctors.h).
template <class R>
PPK_HD double update_cell(const double *__restrict__ Uin, double *__restrict__ Uout,
                          const R *__restrict__ flux, const double *__restrict__ emf,
                          const DevParams &p, long long t, double dtdx, double dtdy, double dtdz,
                          bool fold) {
  const long long vs = p.vstride, sx = 1, sy = p.isize, sz = p.plane;
  const int i = (int)(t % p.isize);
  const long long r = t / p.isize;
  const int j = (int)(r % p.jsize), k = (int)(r / p.jsize);
  const int g = p.gw;
  bool in = i >= g && i < p.isize - g && j >= g && j < p.jsize - g;
  if (p.dim == 3) in = in && k >= g && k < p.ksize - g;
  St<double> u;
#pragma unroll
  for (int v = 0; v < 8; ++v) u[v] = Uin[t + v * vs];
  double inv = 0.0;
  if (in) {
    if (p.dim == 3) {
      auto F = [&](int d, int m, long long off) {
        return double(flux[((long long)d * kFluxComp3 + m) * vs + t + off]);
      };
      u[ID] += F(0, 0, 0) * dtdx; u[IP] += F(0, 1, 0) * dtdx; u[IU] += F(0, 2, 0) * dtdx;
      u[IV] += F(0, 3, 0) * dtdx; u[IW] += F(0, 4, 0) * dtdx;
      u[ID] -= F(0, 0, sx) * dtdx; u[IP] -= F(0, 1, sx) * dtdx; u[IU] -= F(0, 2, sx) * dtdx;
      u[IV] -= F(0, 3, sx) * dtdx; u[IW] -= F(0, 4, sx) * dtdx;
      u[ID] += F(1, 0, 0) * dtdy; u[IP] += F(1, 1, 0) * dtdy; u[IU] += F(1, 3, 0) * dtdy;
      u[IV] += F(1, 2, 0) * dtdy; u[IW] += F(1, 4, 0) * dtdy;
      u[ID] -= F(1, 0, sy) * dtdy; u[IP] -= F(1, 1, sy) * dtdy; u[IU] -= F(1, 3, sy) * dtdy;
      u[IV] -= F(1, 2, sy) * dtdy; u[IW] -= F(1, 4, sy) * dtdy;
      u[ID] += F(2, 0, 0) * dtdz; u[IP] += F(2, 1, 0) * dtdz; u[IU] += F(2, 4, 0) * dtdz;
      u[IV] += F(2, 3, 0) * dtdz; u[IW] += F(2, 2, 0) * dtdz;
      u[ID] -= F(2, 0, sz) * dtdz; u[IP] -= F(2, 1, sz) * dtdz; u[IU] -= F(2, 4, sz) * dtdz;
      u[IV] -= F(2, 3, sz) * dtdz; u[IW] -= F(2, 2, sz) * dtdz;
      u[IA] = ct3_a(u[IA], emf, vs, t, sy, sz, dtdy, dtdz);
      u[IB] = ct3_b(u[IB], emf, vs, t, sx, sz, dtdx, dtdz);
      u[IC] = ct3_c(u[IC], emf, vs, t, sx, sy, dtdx, dtdy);
      if (fold) {
        double mx, my, mz;
        if (i + 1 < p.isize - g) {
          mx = ct3_a(Uin[t + sx + IA * vs], emf, vs, t + sx, sy, sz, dtdy, dtdz);
        } else if (p.bc[1] == PPK_BC_PERIODIC) {
          const long long tw = t + sx - p.nx;
          mx = ct3_a(Uin[tw + IA * vs], emf, vs, tw, sy, sz, dtdy, dtdz);
        } else {
          mx = p.bc[1] == PPK_BC_DIRICHLET ? -u[IA] : u[IA];
        }
        if (j + 1 < p.jsize - g) {
          my = ct3_b(Uin[t + sy + IB * vs], emf, vs, t + sy, sx, sz, dtdx, dtdz);
        } else if (p.bc[3] == PPK_BC_PERIODIC) {
          const long long tw = t + sy - (long long)p.ny * sy;
          my = ct3_b(Uin[tw + IB * vs], emf, vs, tw, sx, sz, dtdx, dtdz);
        } else {
          my = p.bc[3] == PPK_BC_DIRICHLET ? -u[IB] : u[IB];
        }
        if (k + 1 < p.ksize - g) {
          mz = ct3_c(Uin[t + sz + IC * vs], emf, vs, t + sz, sx, sy, dtdx, dtdy);
        } else if (p.bc[5] == PPK_BC_PERIODIC) {
          const long long tw = t + sz - (long long)p.nz * sz;
          mz = ct3_c(Uin[tw + IC * vs], emf, vs, tw, sx, sy, dtdx, dtdy);
        } else {
          mz = p.bc[5] == PPK_BC_DIRICHLET ? -u[IC] : u[IC];
        }
        St<double> q;
        constoprim(u, mx, my, mz, q, p);
        inv = inv_dt_cell(q, p);
      }
    } else {
      auto F = [&](int d, int m, long long off) {
        return double(flux[((long long)d * kFluxComp2 + m) * vs + t + off]);
      };
      u[ID] += F(0, 0, 0) * dtdx; u[IP] += F(0, 1, 0) * dtdx; u[IU] += F(0, 2, 0) * dtdx;
      u[IV] += F(0, 3, 0) * dtdx; u[IW] += F(0, 4, 0) * dtdx; u[IC] += F(0, 5, 0) * dtdx;
      u[ID] -= F(0, 0, sx) * dtdx; u[IP] -= F(0, 1, sx) * dtdx; u[IU] -= F(0, 2, sx) * dtdx;
      u[IV] -= F(0, 3, sx) * dtdx; u[IW] -= F(0, 4, sx) * dtdx; u[IC] -= F(0, 5, sx) * dtdx;
      u[ID] += F(1, 0, 0) * dtdy; u[IP] += F(1, 1, 0) * dtdy; u[IU] += F(1, 3, 0) * dtdy;
      u[IV] += F(1, 2, 0) * dtdy; u[IW] += F(1, 4, 0) * dtdy; u[IC] += F(1, 5, 0) * dtdy;
      u[ID] -= F(1, 0, sy) * dtdy; u[IP] -= F(1, 1, sy) * dtdy; u[IU] -= F(1, 3, sy) * dtdy;
      u[IV] -= F(1, 2, sy) * dtdy; u[IW] -= F(1, 4, sy) * dtdy; u[IC] -= F(1, 5, sy) * dtdy;
      u[IA] = ct2_a(u[IA], emf, t, sy, dtdy);
      u[IB] = ct2_b(u[IB], emf, t, sx, dtdx);
      if (fold) {
        double mx, my;
        if (i + 1 < p.isize - g) {
          mx = ct2_a(Uin[t + sx + IA * vs], emf, t + sx, sy, dtdy);
        } else if (p.bc[1] == PPK_BC_PERIODIC) {
          const long long tw = t + sx - p.nx;
          mx = ct2_a(Uin[tw + IA * vs], emf, tw, sy, dtdy);
        } else {
          mx = p.bc[1] == PPK_BC_DIRICHLET ? -u[IA] : u[IA];
        }
        if (j + 1 < p.jsize - g) {
          my = ct2_b(Uin[t + sy + IB * vs], emf, t + sy, sx, dtdx);
        } else if (p.bc[3] == PPK_BC_PERIODIC) {
          const long long tw = t + sy - (long long)p.ny * sy;
          my = ct2_b(Uin[tw + IB * vs], emf, tw, sx, dtdx);
        } else {
          my = p.bc[3] == PPK_BC_DIRICHLET ? -u[IB] : u[IB];
        }
        St<double> q;
        constoprim(u, mx, my, 0.0, q, p);
        inv = inv_dt_cell(q, p);
      }
    }
  }
#pragma unroll
  for (int v = 0; v < 8; ++v) Uout[t + v * vs] = u[v];
  return inv;
}

// CFL term of padded-box cell t of a state whose ghosts are filled (0 outside the interior),
// as in the staged primitives kernel
PPK_HD double cfl_cell(const double *__restrict__ U, const DevParams &p, long long t) {
  const long long vs = p.vstride;
  const int i = (int)(t % p.isize);
  const long long r = t / p.isize;
  const int j = (int)(r % p.jsize), k = (int)(r / p.jsize);
  const int g = p.gw;
  bool in = i >= g && i < p.isize - g && j >= g && j < p.jsize - g;
  if (p.dim == 3) in = in && k >= g && k < p.ksize - g;
  if (!in) return 0.0;
  St<double> u, q;
#pragma unroll
  for (int v = 0; v < 8; ++v) u[v] = U[t + v * vs];
  const double mx = U[t + 1 + IA * vs];
  const double my = U[t + p.isize + IB * vs];
  const double mz = p.dim == 3 ? U[t + p.plane + IC * vs] : 0.0;
  constoprim(u, mx, my, mz, q, p);
  return inv_dt_cell(q, p);
}

}  // namespace fusedk
}  // namespace PPK_NS
}  // namespace ppk
