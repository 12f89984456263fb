// Single-pass 3D MHD step (k_march3d): per-thread bodies, host+device.
//
// A block owns an 18 x 14 tile of record columns (x, y) and marches along z through a range of
// planes (a block's share of the grid may span several tiles: MarchGeom).  One thread per
// column (256 threads, the last 4 without a column).  At plane r the block
//   H  computes the half-step record of every column at plane r (primitives, edge electric
//      field, limited slopes, Hancock predictor; the same arithmetic as k_half3d) and puts it in
//      LDS,
//   S  solves, per column c: the x and y faces and the z edge of plane r (records of plane r
//      from LDS), and the z face and the x and y edges between planes r-1 and r (their plane
//      r-1 states carried in registers from the previous plane),
//   U  finishes the update + CT of plane r-1 and starts that of plane r (the reference's order
//      +F(i) -F(i+1) per direction and its CT order, one partial sum kept across the plane),
//      writes plane r-1 of the new state and folds its CFL term into a per-block maximum.
// The interior 16 x 12 columns are the block's output (16 x 12 tiles cover a 256-wide grid
// exactly along x and waste 3 % along y, where 14 x 14 tiles wasted 7 % of the work at 256^3);
// the one-column ring around them
// provides the records that the output cells' faces and edges need.  Each record field, each
// face flux and each EMF of the output is computed once per block (faces and edges at the
// ring are computed by both neighbouring blocks).
//
// The upper-side face values of a cell (AR, BR, CR and their transverse slopes) are equal,
// bit for bit, to the lower-side values of its +1 neighbour (same expression on the same
// operands).  The faces and edges of plane r take AR / BR from the x / y neighbour's record in
// LDS, as the three-kernel path does; CR and its slopes (needed by the z face and the x / y
// edges between planes r and r+1, solved one plane later) are computed by the cell itself, so
// only one plane of records is held.
//
// LDS (doubles): two plane slots, each Q (8 x 20 x 16, origin x0-1), face fields (3 x 21 x 17,
// origin x0-1) and E (3 x 19 x 15, origin x0), and the records of plane r (41 fields at a
// stride of 257): 156,072 bytes, one block per CU.  Slot r & 1 holds plane r; after H(r) it is free and serves
// as the exchange area of S/U, then receives plane r + 2.
//
// The arithmetic is that of the three-kernel fused path and of the staged kernels, so the
// strict build is bit-identical to the reference's double arithmetic.  References: the half
// step src/muscl/MHDBaseFunctor3D.h:688-1114, faces src/muscl/MHDRunFunctors3D.h:1839-1902,
// EMFs :2183-2230, update :2430-2544, CT :2549-2628.
#pragma once

#include "fused_body.hpp"

namespace ppk {
namespace PPK_NS {
namespace marchk {

using namespace fusedk;

constexpr int kMTX = 18, kMTY = 14;      // record tile (columns along x, y)
constexpr int kMOX = kMTX - 2, kMOY = kMTY - 2;  // output tile
constexpr int kMC = kMTX * kMTY;         // record columns (threads kMC .. kMN-1 have none)
constexpr int kMN = 256;                 // threads per block; also the stride of the per-column
                                         // LDS arrays (records, exchange)
static_assert(kMC <= kMN, "one thread per record column");
constexpr int kQW = kMTX + 2, kFW = kMTX + 3, kEW = kMTX + 1;  // row widths of Q, faces, E
constexpr int kQN = kQW * (kMTY + 2), kFN = kFW * (kMTY + 3), kEN = kEW * (kMTY + 1);
static_assert(kQN <= 2 * kMN && kFN <= 2 * kMN && kEN <= 2 * kMN, "two cells per thread");
constexpr int kSlotF = 8 * kQN, kSlotE = kSlotF + 3 * kFN, kSlot = kSlotE + 3 * kEN;
// record fields in LDS: the 38 of a three-kernel record, then the cell's own CR and its
// transverse slopes
enum { R3_CR = kRec3, R3_DCRX, R3_DCRY, R3X_N };
constexpr int kRecOff = 2 * kSlot;
// record field stride: one double more than the columns' stride, so that the fields of a
// column fall on different LDS banks
constexpr int kRS = kMN + 1;
constexpr int kMarchLds = kRecOff + R3X_N * kRS;
static_assert(kMarchLds * 8 <= 160 * 1024, "march tile exceeds the LDS of a CU");
// exchange area of S -> U in the free slot: x flux (5), y flux (5), Ez, Ex, Ey per column
enum { X_FX = 0, X_FY = 5, X_EZ = 10, X_EX = 11, X_EY = 12, X_N = 13 };
static_assert(X_N * kMN <= kSlotE, "exchange area");
// U -> CFL: the new A and B of plane r-1, in the E part of the free slot
static_assert(2 * kMN <= 3 * kEN, "CFL exchange area");

// block -> work, with one block per CU (tiles x fastest, then y):
//  - rounds of whole tiles: in round k block L marches every plane of tile k * nblk + L, so
//    the blocks of one XCD (consecutive L) march neighbouring tiles side by side and share
//    their halo columns in that XCD's L2;
//  - the tiles left over (fewer than nblk) are shared out: their tile-planes in tile order
//    (every plane of a tile, then the next tile) are cut into nblk ranges of `per`, and block L
//    marches range L, one segment per tile it touches.
// Every CU gets the same number of planes (a launch of fixed-length z chunks ends with a
// partial round of blocks), at the cost of a two-plane start per segment.
struct MarchGeom {
  int ntx, nty;  // tiles along x and y
  int nblk;      // blocks
  int full;      // rounds of whole tiles
  int per;       // tile-planes per block of the left-over tiles (the last ranges may be
                 // shorter or empty)
  long long blocks() const { return nblk; }
};
inline MarchGeom march_geom(const DevParams &p, int nblk) {
  MarchGeom g;
  g.ntx = (p.nx + kMOX - 1) / kMOX;
  g.nty = (p.ny + kMOY - 1) / kMOY;
  const long long total = (long long)g.ntx * g.nty * p.nz;
  long long b = nblk < 1 ? 1 : nblk;
  if (b > total) b = total;
  g.nblk = (int)b;
  g.full = g.ntx * g.nty / g.nblk;
  const long long rest = total - (long long)g.full * g.nblk * p.nz;
  g.per = (int)((rest + b - 1) / b);
  return g;
}
// tile t: record columns x0 .. x0+17 (output x0+1 .. x0+16), y0 .. y0+13 (output y0+1 ..
// y0+12)
PPK_HD void march_tile(const DevParams &p, const MarchGeom &g, int t, int &x0, int &y0) {
  x0 = p.gw - 1 + kMOX * (t % g.ntx);
  y0 = p.gw - 1 + kMOY * (t / g.ntx);
}

PPK_HD double *slot(double *lds, int s) { return lds + s * kSlot; }

// a record of plane r from LDS (column n of the tile); upper-side A / B values zero, as
// load_rec3 leaves them
PPK_HD void lds_rec(const double *sR, int n, Half3D<double> &h) {
  const double *a = sR + n;
  constexpr int S = kRS;
  h.r = a[R3_R * S];     h.pp = a[R3_P * S];    h.u = a[R3_U * S];     h.v = a[R3_V * S];
  h.w = a[R3_W * S];     h.A = a[R3_A * S];     h.B = a[R3_B * S];     h.C = a[R3_C * S];
  h.drx = a[R3_DRX * S]; h.dpx = a[R3_DPX * S]; h.dux = a[R3_DUX * S]; h.dvx = a[R3_DVX * S];
  h.dwx = a[R3_DWX * S]; h.dBx = a[R3_DBX * S]; h.dCx = a[R3_DCX * S];
  h.dry = a[R3_DRY * S]; h.dpy = a[R3_DPY * S]; h.duy = a[R3_DUY * S]; h.dvy = a[R3_DVY * S];
  h.dwy = a[R3_DWY * S]; h.dAy = a[R3_DAY * S]; h.dCy = a[R3_DCY * S];
  h.drz = a[R3_DRZ * S]; h.dpz = a[R3_DPZ * S]; h.duz = a[R3_DUZ * S]; h.dvz = a[R3_DVZ * S];
  h.dwz = a[R3_DWZ * S]; h.dAz = a[R3_DAZ * S]; h.dBz = a[R3_DBZ * S];
  h.AL = a[R3_AL * S];   h.BL = a[R3_BL * S];   h.CL = a[R3_CL * S];
  h.dALy = a[R3_DALY * S]; h.dALz = a[R3_DALZ * S]; h.dBLx = a[R3_DBLX * S];
  h.dBLz = a[R3_DBLZ * S]; h.dCLx = a[R3_DCLX * S]; h.dCLy = a[R3_DCLY * S];
  h.CR = a[R3_CR * S];   h.dCRx = a[R3_DCRX * S]; h.dCRy = a[R3_DCRY * S];
  h.AR = h.BR = 0.0;
  h.dARy = h.dARz = h.dBRx = h.dBRz = 0.0;
}
PPK_HD void lds_store_rec(double *sR, int n, const Half3D<double> &h) {
  double *o = sR + n;
  constexpr int S = kRS;
  o[R3_R * S] = h.r;     o[R3_P * S] = h.pp;    o[R3_U * S] = h.u;     o[R3_V * S] = h.v;
  o[R3_W * S] = h.w;     o[R3_A * S] = h.A;     o[R3_B * S] = h.B;     o[R3_C * S] = h.C;
  o[R3_DRX * S] = h.drx; o[R3_DPX * S] = h.dpx; o[R3_DUX * S] = h.dux; o[R3_DVX * S] = h.dvx;
  o[R3_DWX * S] = h.dwx; o[R3_DBX * S] = h.dBx; o[R3_DCX * S] = h.dCx;
  o[R3_DRY * S] = h.dry; o[R3_DPY * S] = h.dpy; o[R3_DUY * S] = h.duy; o[R3_DVY * S] = h.dvy;
  o[R3_DWY * S] = h.dwy; o[R3_DAY * S] = h.dAy; o[R3_DCY * S] = h.dCy;
  o[R3_DRZ * S] = h.drz; o[R3_DPZ * S] = h.dpz; o[R3_DUZ * S] = h.duz; o[R3_DVZ * S] = h.dvz;
  o[R3_DWZ * S] = h.dwz; o[R3_DAZ * S] = h.dAz; o[R3_DBZ * S] = h.dBz;
  o[R3_AL * S] = h.AL;   o[R3_BL * S] = h.BL;   o[R3_CL * S] = h.CL;
  o[R3_DALY * S] = h.dALy; o[R3_DALZ * S] = h.dALz; o[R3_DBLX * S] = h.dBLx;
  o[R3_DBLZ * S] = h.dBLz; o[R3_DCLX * S] = h.dCLx; o[R3_DCLY * S] = h.dCLy;
  o[R3_CR * S] = h.CR;   o[R3_DCRX * S] = h.dCRx; o[R3_DCRY * S] = h.dCRy;
}

template <class R>
PPK_HD void swap2(St<R> &s, int a, int b) {
  R t = s[a];
  s[a] = s[b];
  s[b] = t;
}

// The registers of one thread (column) of a marching block.
template <int RS, int ST = 0>
struct MarchThread {
  int tid, a, b;     // thread, column position in the tile
  int i, j;          // column's cell indices
  bool rok;          // the column has records (i, j <= size - gw)
  bool ocol;         // output column: 1 <= a, b <= 14 and an interior cell
  bool acol, bcol;   // its new A (x-neighbour's CFL) / new B (y-neighbour's CFL) is needed
  // plane r-1 values of the own column for the half step of plane r
  St<double> Qm;
  double Am, Bm;
  // states of plane r-1 for the solves between planes r-1 and r: upper z face of c, and the
  // upper (+z) corner states of the x and y edges: LT of c, RT of the -y (x edge) / -x (y
  // edge) neighbour
  St<double> qmz, ltx, rtx, lty, rty;
  // update of plane r-1 in progress: hydro sums through +Fz(r-1), CT partials of A and B, the
  // final C, and the y / x EMFs of the lower edges (for the z differences)
  double up[5], Ap, Bp, Cf, Eym, Exm;
  St<double> Uc;     // state of plane r (the update's input)
  St<double> un;     // new state of plane r-1 (for its CFL term)
  double Cn;         // new C of plane r (the CFL term of plane r-1 needs it as mz)
  // results of S(r) for U(r)
  double fx[5], fy[5], fz[5], ex, ey, ez;
  Half3D<double> h;  // the record of plane r
  // prefetched loads of plane r+2: Q cells (8 vars, A(i+1), B(j+1), C(k+1)) and face cells
  double pq[2][11], pf[2][3];
  bool pqok[2], pfok[2];
  double cfl;        // maximum CFL term of this thread's output cells (over every segment)

  PPK_HD void init(const DevParams &p, int t, int x0, int y0) {
    tid = t;
    a = t % kMTX;
    b = t / kMTX;
    i = x0 + a;
    j = y0 + b;
    rok = t < kMC && i <= p.isize - p.gw && j <= p.jsize - p.gw;
    const bool in = i < p.isize - p.gw && j < p.jsize - p.gw;
    ocol = t < kMC && a >= 1 && a <= kMOX && b >= 1 && b <= kMOY && in;
    acol = rok && a >= 1 && b >= 1 && b <= kMOY && j < p.jsize - p.gw;
    bcol = rok && b >= 1 && a >= 1 && a <= kMOX && i < p.isize - p.gw;
  }

  // global loads of plane pl for its Q and face cells (issued early, stashed later)
  PPK_HD void prefetch(const double *__restrict__ U, const DevParams &p, int x0, int y0,
                       int pl) {
    const long long vs = p.vstride;
    const int qhi = 1 - p.gw;  // Q cells up to size + qhi (the reference's primitive range)
#pragma unroll
    for (int it = 0; it < 2; ++it) {
      const int n = tid + it * kMN;
      const int ii = x0 - 1 + n % kQW, jj = y0 - 1 + n / kQW;
      pqok[it] = n < kQN && ii <= p.isize + qhi && jj <= p.jsize + qhi;
      const long long t = cidx(p, min(ii, p.isize + qhi), min(jj, p.jsize + qhi), pl);
#pragma unroll
      for (int v = 0; v < 8; ++v) pq[it][v] = U[t + v * vs];
      pq[it][8] = U[t + 1 + IA * vs];
      pq[it][9] = U[t + p.isize + IB * vs];
      pq[it][10] = U[t + p.plane + IC * vs];
    }
#pragma unroll
    for (int it = 0; it < 2; ++it) {
      const int n = tid + it * kMN;
      const int ii = x0 - 1 + n % kFW, jj = y0 - 1 + n / kFW;
      pfok[it] = n < kFN && ii < p.isize && jj < p.jsize;
      const long long t = cidx(p, min(ii, p.isize - 1), min(jj, p.jsize - 1), pl);
      pf[it][0] = U[t + IA * vs];
      pf[it][1] = U[t + IB * vs];
      pf[it][2] = U[t + IC * vs];
    }
  }

  // primitives of the prefetched Q cells (MHDRunFunctors3D.h:88-163) and the face fields,
  // into slot s
  PPK_HD void stash(const DevParams &p, double *lds, int s) const {
    double *sQ = slot(lds, s), *sF = sQ + kSlotF;
#pragma unroll
    for (int it = 0; it < 2; ++it) {
      if (!pqok[it]) continue;
      const int n = tid + it * kMN;
      St<double> u, q;
#pragma unroll
      for (int v = 0; v < 8; ++v) u[v] = pq[it][v];
      constoprim(u, pq[it][8], pq[it][9], pq[it][10], q, p);
#pragma unroll
      for (int v = 0; v < 8; ++v) sQ[v * kQN + n] = q[v];
    }
#pragma unroll
    for (int it = 0; it < 2; ++it) {
      if (!pfok[it]) continue;
      const int n = tid + it * kMN;
      sF[n] = pf[it][0];
      sF[kFN + n] = pf[it][1];
      sF[2 * kFN + n] = pf[it][2];
    }
  }

  // own Q and face fields of the plane in slot s, as the half step of the next plane needs them
  PPK_HD void keep_lower(const double *lds, int s) {
    const double *sQ = lds + s * kSlot, *sF = sQ + kSlotF;
    const int nq = (a + 1) + kQW * (b + 1), nf = (a + 1) + kFW * (b + 1);
#pragma unroll
    for (int v = 0; v < 8; ++v) Qm[v] = sQ[v * kQN + nq];
    Am = sF[nf];
    Bm = sF[kFN + nf];
  }

  // edge electric field of plane pl (slot s) from Q and faces of planes pl-1 (slot sm) and pl,
  // same operations as half3_elec (MHDRunFunctors3D.h:303-356)
  PPK_HD void elec(const DevParams &p, double *lds, int s, int sm, int x0, int y0) const {
    const double *qa = slot(lds, sm), *qb = slot(lds, s);
    const double *fa = qa + kSlotF, *fb = qb + kSlotF;
    double *sE = slot(lds, s) + kSlotE;
    const int hi = 1 - p.gw;
#pragma unroll
    for (int it = 0; it < 2; ++it) {
      const int n = tid + it * kMN;
      if (n >= kEN) continue;
      const int ex = n % kEW, ey = n / kEW;
      if (x0 + ex > p.isize + hi || y0 + ey > p.jsize + hi) continue;
      const int X = ex + 1, Y = ey + 1;  // position in the Q / face regions
      auto q = [&](int dx, int dy, int dz, int v) {
        return (dz < 0 ? qa : qb)[v * kQN + (X + dx) + kQW * (Y + dy)];
      };
      auto u = [&](int dx, int dy, int dz, int c) {
        return (dz < 0 ? fa : fb)[c * kFN + (X + dx) + kFW * (Y + dy)];
      };
      double v_, w_, u_, A, B, C;
      v_ = 0.25 * (q(0, -1, -1, IV) + q(0, -1, 0, IV) + q(0, 0, -1, IV) + q(0, 0, 0, IV));
      w_ = 0.25 * (q(0, -1, -1, IW) + q(0, -1, 0, IW) + q(0, 0, -1, IW) + q(0, 0, 0, IW));
      B = 0.5 * (u(0, 0, -1, 1) + u(0, 0, 0, 1));
      C = 0.5 * (u(0, -1, 0, 2) + u(0, 0, 0, 2));
      sE[n] = v_ * C - w_ * B;
      u_ = 0.25 * (q(-1, 0, -1, IU) + q(-1, 0, 0, IU) + q(0, 0, -1, IU) + q(0, 0, 0, IU));
      w_ = 0.25 * (q(-1, 0, -1, IW) + q(-1, 0, 0, IW) + q(0, 0, -1, IW) + q(0, 0, 0, IW));
      A = 0.5 * (u(0, 0, -1, 0) + u(0, 0, 0, 0));
      C = 0.5 * (u(-1, 0, 0, 2) + u(0, 0, 0, 2));
      sE[kEN + n] = w_ * A - u_ * C;
      u_ = 0.25 * (q(-1, -1, 0, IU) + q(-1, 0, 0, IU) + q(0, -1, 0, IU) + q(0, 0, 0, IU));
      v_ = 0.25 * (q(-1, -1, 0, IV) + q(-1, 0, 0, IV) + q(0, -1, 0, IV) + q(0, 0, 0, IV));
      A = 0.5 * (u(0, -1, 0, 0) + u(0, 0, 0, 0));
      B = 0.5 * (u(-1, 0, 0, 1) + u(0, 0, 0, 1));
      sE[2 * kEN + n] = u_ * B - v_ * A;
    }
  }

  // the state of plane r that the update starts from (issued early)
  PPK_HD void load_state(const double *__restrict__ U, const DevParams &p, int r) {
    if (!(rok && a >= 1 && b >= 1)) return;
    const long long t = cidx(p, i, j, r);
#pragma unroll
    for (int v = 0; v < 8; ++v) Uc[v] = U[t + v * p.vstride];
  }

  // H(r): the record of the column at plane r (slot sa: plane r, slot sb: plane r+1), same
  // operations as half3_record, plus the cell's own CR and its x / y slopes; into LDS
  PPK_HD void half(const DevParams &p, double *lds, int sa, int sb, double dt) {
    if (!rok) return;
    const double *Qa = slot(lds, sa), *Qb = slot(lds, sb);
    const double *Fa = Qa + kSlotF, *Fb = Qb + kSlotF;
    const double *Ea = Qa + kSlotE, *Eb = Qb + kSlotE;
    const int X = a + 1, Y = b + 1;
    auto ldq = [&](const double *Qs, int dx, int dy, St<double> &s) {
      const int n = (X + dx) + kQW * (Y + dy);
#pragma unroll
      for (int v = 0; v < 8; ++v) s[v] = Qs[v * kQN + n];
    };
    St<double> q, qPX, qMX, qPY, qMY, qPZ;
    ldq(Qa, 0, 0, q);
    ldq(Qa, 1, 0, qPX);
    ldq(Qa, -1, 0, qMX);
    ldq(Qa, 0, 1, qPY);
    ldq(Qa, 0, -1, qMY);
    ldq(Qb, 0, 0, qPZ);
    auto F = [&](const double *Fs, int dx, int dy, int c) {
      return Fs[c * kFN + (X + dx) + kFW * (Y + dy)];
    };
    const double bfNb[6] = {F(Fa, 0, 0, 0), F(Fa, 1, 0, 0), F(Fa, 0, 0, 1),
                            F(Fa, 0, 1, 1), F(Fa, 0, 0, 2), F(Fb, 0, 0, 2)};
    const double xst = ST ? double(ST) : fmin(p.slope_type, 2.0);
    // the left-side transverse slopes, and the cell's own CR slopes (entries 10, 11)
    const double d12[12] = {limited_slope(xst, F(Fa, 0, 0, 0), F(Fa, 0, 1, 0), F(Fa, 0, -1, 0)),
                            limited_slope(xst, F(Fa, 0, 0, 0), F(Fb, 0, 0, 0), Am),
                            limited_slope(xst, F(Fa, 0, 0, 1), F(Fa, 1, 0, 1), F(Fa, -1, 0, 1)),
                            limited_slope(xst, F(Fa, 0, 0, 1), F(Fb, 0, 0, 1), Bm),
                            limited_slope(xst, F(Fa, 0, 0, 2), F(Fa, 1, 0, 2), F(Fa, -1, 0, 2)),
                            limited_slope(xst, F(Fa, 0, 0, 2), F(Fa, 0, 1, 2), F(Fa, 0, -1, 2)),
                            0.0, 0.0, 0.0, 0.0,
                            limited_slope(xst, F(Fb, 0, 0, 2), F(Fb, 1, 0, 2), F(Fb, -1, 0, 2)),
                            limited_slope(xst, F(Fb, 0, 0, 2), F(Fb, 0, 1, 2), F(Fb, 0, -1, 2))};
    auto E = [&](const double *Es, int dx, int dy, int c) {
      return Es[c * kEN + (a + dx) + kEW * (b + dy)];
    };
    const double Ex[4] = {E(Ea, 0, 0, 0), E(Eb, 0, 0, 0), E(Ea, 0, 1, 0), E(Eb, 0, 1, 0)};
    const double Ey[4] = {E(Ea, 0, 0, 1), E(Eb, 0, 0, 1), E(Ea, 1, 0, 1), E(Eb, 1, 0, 1)};
    const double Ez[4] = {E(Ea, 0, 0, 2), E(Ea, 0, 1, 2), E(Ea, 1, 0, 2), E(Ea, 1, 1, 2)};
    const double dtdx = dt / p.dx, dtdy = dt / p.dy, dtdz = dt / p.dz;
    half_step_3d<ST>(q, qPX, qMX, qPY, qMY, qPZ, Qm, bfNb, d12, Ex, Ey, Ez, dtdx, dtdy, dtdz, p, h);
    lds_store_rec(lds + kRecOff, tid, h);
    Qm = q;
    Am = bfNb[0];
    Bm = bfNb[2];
  }

  // S(r): the solves of the column.  mid: plane r is an output plane; cross: the z face and
  // the x / y edges between planes r-1 and r are needed; carry: the plane-r states of the
  // solves between planes r and r+1 are needed.  pre() (the global loads of a later plane,
  // which every thread issues) runs between the face and the edge solves, so that its
  // registers are not held through the face solves.
  template <class Pre>
  PPK_HD void solves(const DevParams &p, double *lds, int sa, bool mid, bool cross, bool carry,
                     Pre &&pre) {
    if (!rok) {
      pre();
      return;
    }
    const double *sR = lds + kRecOff;
    const bool inA = a >= 1, inB = b >= 1;
    if (mid && inA && inB && b <= kMOY) {  // x face of c (its lower face)
      Half3D<double> hl;
      lds_rec(sR, tid - 1, hl);
      hl.AR = h.AL;
      St<double> ql = face_state_3d(hl, 0, +1, p), qr = face_state_3d(h, 0, -1, p), f;
      riemann_mhd_rs<RS>(ql, qr, f, p);
      fx[0] = f[ID]; fx[1] = f[IP]; fx[2] = f[IU]; fx[3] = f[IV]; fx[4] = f[IW];
    }
    if (mid && inA && inB && a <= kMOX) {  // y face
      Half3D<double> hl;
      lds_rec(sR, tid - kMTX, hl);
      hl.BR = h.BL;
      St<double> ql = face_state_3d(hl, 1, +1, p), qr = face_state_3d(h, 1, -1, p), f;
      swap2(ql, IU, IV); swap2(ql, IA, IB);
      swap2(qr, IU, IV); swap2(qr, IA, IB);
      riemann_mhd_rs<RS>(ql, qr, f, p);
      fy[0] = f[ID]; fy[1] = f[IP]; fy[2] = f[IU]; fy[3] = f[IV]; fy[4] = f[IW];
    }
    if (cross && inA && inB && a <= kMOX && b <= kMOY) {  // z face between r-1 and r
      St<double> ql = qmz, qr = face_state_3d(h, 2, -1, p), f;
      swap2(ql, IU, IW); swap2(ql, IA, IC);
      swap2(qr, IU, IW); swap2(qr, IA, IC);
      riemann_mhd_rs<RS>(ql, qr, f, p);
      fz[0] = f[ID]; fz[1] = f[IP]; fz[2] = f[IU]; fz[3] = f[IV]; fz[4] = f[IW];
    }
    pre();
    if (mid && inA && inB) {  // z edge: RT of c-x-y, RB of c-x, LT of c-y, LB of c
      Half3D<double> ra, rb, rd, t;
      lds_rec(sR, tid - 1 - kMTX, ra);
      lds_rec(sR, tid - kMTX, t);
      ra.AR = t.AL; ra.dARy = t.dALy;
      lds_rec(sR, tid - 1, rb);
      ra.BR = rb.BL; ra.dBRx = rb.dBLx;
      rb.AR = h.AL; rb.dARy = h.dALy;
      rd = t;
      rd.BR = h.BL; rd.dBRx = h.dBLx;
      ez = compute_emf<EMFZ>(edge_state_3d(ra, 2, IRT, p), edge_state_3d(rb, 2, IRB, p),
                             edge_state_3d(rd, 2, ILT, p), edge_state_3d(h, 2, ILB, p), p);
    }
    if (cross && inA && inB && a <= kMOX) {  // x edge: RT of c-y-z, RB of c-y, LT of c-z, LB of c
      Half3D<double> rb;
      lds_rec(sR, tid - kMTX, rb);
      rb.BR = h.BL; rb.dBRz = h.dBLz;
      ex = compute_emf<EMFX>(rtx, edge_state_3d(rb, 0, IRB, p), ltx, edge_state_3d(h, 0, ILB, p),
                             p);
    }
    if (cross && inA && inB && b <= kMOY) {  // y edge: RT of c-x-z, LT of c-z, RB of c-x, LB of c
      Half3D<double> rd;
      lds_rec(sR, tid - 1, rd);
      rd.AR = h.AL; rd.dARz = h.dALz;
      ey = compute_emf<EMFY>(rty, lty, edge_state_3d(rd, 1, IRB, p), edge_state_3d(h, 1, ILB, p),
                             p);
    }
    if (carry) {  // plane-r states of the solves between planes r and r+1
      if (inA && inB && a <= kMOX && b <= kMOY) qmz = face_state_3d(h, 2, +1, p);
      if (inA && inB && a <= kMOX) {
        ltx = edge_state_3d(h, 0, ILT, p);
        Half3D<double> ra;
        lds_rec(sR, tid - kMTX, ra);
        ra.BR = h.BL; ra.dBRz = h.dBLz;
        rtx = edge_state_3d(ra, 0, IRT, p);
      }
      if (inA && inB && b <= kMOY) {
        lty = edge_state_3d(h, 1, ILT, p);
        Half3D<double> ra;
        lds_rec(sR, tid - 1, ra);
        ra.AR = h.AL; ra.dARz = h.dALz;
        rty = edge_state_3d(ra, 1, IRT, p);
      }
    }
    // exchange for the update (slot sa is free since H(r))
    double *sX = slot(lds, sa);
#pragma unroll
    for (int m = 0; m < 5; ++m) {
      sX[(X_FX + m) * kMN + tid] = fx[m];
      sX[(X_FY + m) * kMN + tid] = fy[m];
    }
    sX[X_EZ * kMN + tid] = ez;
    sX[X_EX * kMN + tid] = ex;
    sX[X_EY * kMN + tid] = ey;
  }

  // U(r): finish plane r-1 (done: r-1 is an output plane) and start plane r (mid), or only
  // the new C of plane r (top: the plane above the last output plane, for its CFL term)
  PPK_HD void update(const DevParams &p, double *lds, int sa, double dt, bool done, bool mid,
                     bool top, double *__restrict__ Uout, int r, bool fold) {
    if (!rok || a < 1 || b < 1) return;
    const double dtdx = dt / p.dx, dtdy = dt / p.dy, dtdz = dt / p.dz;
    const double *sX = slot(lds, sa);
    double *sC = slot(lds, sa) + kSlotE;
    auto X = [&](int f, int n) { return sX[f * kMN + n]; };
    if (done) {
      if (ocol) {
        un[ID] = up[0] - fz[0] * dtdz;
        un[IP] = up[1] - fz[1] * dtdz;
        un[IU] = up[2] - fz[4] * dtdz;
        un[IV] = up[3] - fz[3] * dtdz;
        un[IW] = up[4] - fz[2] * dtdz;
        un[IA] = Ap - (ey - Eym) * dtdz;
        un[IB] = Bp + (ex - Exm) * dtdz;
        un[IC] = Cf;
        const long long t = cidx(p, i, j, r - 1);
#pragma unroll
        for (int v = 0; v < 8; ++v) Uout[t + v * p.vstride] = un[v];
      }
      if (fold) {
        if (acol) sC[tid] = Ap - (ey - Eym) * dtdz;
        if (bcol) sC[kMN + tid] = Bp + (ex - Exm) * dtdz;
      }
    }
    if (mid) {
      if (ocol) {
        double u[5] = {Uc[ID], Uc[IP], Uc[IU], Uc[IV], Uc[IW]};
        const int nx = tid + 1, ny = tid + kMTX;
        // x: +F(i) -F(i+1); y: +F(j) -F(j+1) (y and z fluxes stored rotated); z: +F(k)
        u[0] += fx[0] * dtdx; u[1] += fx[1] * dtdx; u[2] += fx[2] * dtdx;
        u[3] += fx[3] * dtdx; u[4] += fx[4] * dtdx;
        u[0] -= X(X_FX + 0, nx) * dtdx; u[1] -= X(X_FX + 1, nx) * dtdx;
        u[2] -= X(X_FX + 2, nx) * dtdx; u[3] -= X(X_FX + 3, nx) * dtdx;
        u[4] -= X(X_FX + 4, nx) * dtdx;
        u[0] += fy[0] * dtdy; u[1] += fy[1] * dtdy; u[2] += fy[3] * dtdy;
        u[3] += fy[2] * dtdy; u[4] += fy[4] * dtdy;
        u[0] -= X(X_FY + 0, ny) * dtdy; u[1] -= X(X_FY + 1, ny) * dtdy;
        u[2] -= X(X_FY + 3, ny) * dtdy; u[3] -= X(X_FY + 2, ny) * dtdy;
        u[4] -= X(X_FY + 4, ny) * dtdy;
        u[0] += fz[0] * dtdz; u[1] += fz[1] * dtdz; u[2] += fz[4] * dtdz;
        u[3] += fz[3] * dtdz; u[4] += fz[2] * dtdz;
#pragma unroll
        for (int m = 0; m < 5; ++m) up[m] = u[m];
      }
      if (ocol || acol) {  // A += (Ez(j+1) - Ez(j)) dt/dy, then later -= (Ey(k+1) - Ey(k)) dt/dz
        double A = Uc[IA];
        A += (X(X_EZ, tid + kMTX) - ez) * dtdy;
        Ap = A;
        Eym = ey;
      }
      if (ocol || bcol) {  // B -= (Ez(i+1) - Ez(i)) dt/dx, then later += (Ex(k+1) - Ex(k)) dt/dz
        double B = Uc[IB];
        B -= (X(X_EZ, tid + 1) - ez) * dtdx;
        Bp = B;
        Exm = ex;
      }
    }
    if ((mid || top) && ocol) {  // C += (Ey(i+1) - Ey(i)) dt/dx; C -= (Ex(j+1) - Ex(j)) dt/dy
      double C = Uc[IC];
      C += (X(X_EY, tid + 1) - ey) * dtdx;
      C -= (X(X_EX, tid + kMTX) - ex) * dtdy;
      if (mid) Cf = C;
      Cn = C;
    }
  }

  // CFL term of the new cell of plane r-1 (after the new A / B of the neighbours are in LDS):
  // its primitive needs the new face fields of its +1 neighbours, or past the last interior
  // cell the values the next ghost fill puts there (fusedk::update_cell)
  PPK_HD double cfl_term(const DevParams &p, const double *lds, int sa, int r) {
    if (!ocol) return 0.0;
    const double *sC = lds + sa * kSlot + kSlotE;
    const int g = p.gw;
    double mx, my, mz;
    if (i + 1 < p.isize - g || p.bc[1] == PPK_BC_PERIODIC)
      mx = sC[tid + 1];
    else
      mx = p.bc[1] == PPK_BC_DIRICHLET ? -un[IA] : un[IA];
    if (j + 1 < p.jsize - g || p.bc[3] == PPK_BC_PERIODIC)
      my = sC[kMN + tid + kMTX];
    else
      my = p.bc[3] == PPK_BC_DIRICHLET ? -un[IB] : un[IB];
    if (r < p.ksize - g || p.bc[5] == PPK_BC_PERIODIC)  // r: the plane above, r - 1 + 1
      mz = Cn;
    else
      mz = p.bc[5] == PPK_BC_DIRICHLET ? -un[IC] : un[IC];
    St<double> q;
    constoprim(un, mx, my, mz, q, p);
    const double term = inv_dt_cell(q, p);
    cfl = fmax(cfl, term);
    return term;
  }
};

// One segment of a block's work: output planes [z0, z1) of tile `tile`.  X.each(f) runs
// f(thread) for every thread of the block and then synchronizes the block;
// X.cell_cfl(thread, plane, term) sees each output cell's CFL term (a no-op in the kernel).
template <int RS, int ST, class Exec>
PPK_HD void march_segment(Exec &X, const double *__restrict__ Uin, double *__restrict__ Uout,
                          const DevParams &p, const MarchGeom &g, int tile, int z0, int z1,
                          double dt, bool fold, double *lds) {
  int x0, y0;
  march_tile(p, g, tile, x0, y0);
  using T = MarchThread<RS, ST>;
  X.each([&](T &t) {
    t.init(p, X.tid(t), x0, y0);
    t.prefetch(Uin, p, x0, y0, z0 - 2);
    t.stash(p, lds, (z0 - 2) & 1);
  });
  X.each([&](T &t) {
    t.prefetch(Uin, p, x0, y0, z0 - 1);
    t.keep_lower(lds, (z0 - 2) & 1);
    t.stash(p, lds, (z0 - 1) & 1);
  });
  X.each([&](T &t) { t.elec(p, lds, (z0 - 1) & 1, (z0 - 2) & 1, x0, y0); });
  X.each([&](T &t) {
    t.prefetch(Uin, p, x0, y0, z0);
    t.stash(p, lds, z0 & 1);
  });
  X.each([&](T &t) { t.elec(p, lds, z0 & 1, (z0 - 1) & 1, x0, y0); });
  for (int r = z0 - 1; r <= z1; ++r) {
    const int sa = r & 1, sb = (r + 1) & 1;
    const bool ahead = r + 2 <= z1 + 1;  // plane r+2 is needed (half step of plane r+1)
    const bool mid = r >= z0 && r < z1, cross = r >= z0, carry = r < z1;
    const bool done = r > z0, top = r == z1;
    X.each([&](T &t) {
      if (r >= z0) t.load_state(Uin, p, r);
      t.half(p, lds, sa, sb, dt);
    });
    X.each([&](T &t) {
      t.solves(p, lds, sa, mid, cross, carry, [&] {
        if (ahead) t.prefetch(Uin, p, x0, y0, r + 2);
      });
    });
    X.each([&](T &t) { t.update(p, lds, sa, dt, done, mid, top, Uout, r, fold); });
    X.each([&](T &t) {
      if (fold && done && t.ocol) X.cell_cfl(t, r - 1, t.cfl_term(p, lds, sa, r));
      if (ahead) t.stash(p, lds, sa);
    });
    if (ahead) X.each([&](T &t) { t.elec(p, lds, sa, sb, x0, y0); });
  }
}

// Segment k of block L (MarchGeom): output planes [z0, z1) of tile `tile`; false past the
// last.  32-bit arithmetic (a grid of 2^31 tile-planes would not fit a GPU's memory).
PPK_HD bool march_piece(const DevParams &p, const MarchGeom &g, int L, int k, int &tile, int &z0,
                        int &z1) {
  if (k < g.full) {  // whole tile
    tile = k * g.nblk + L;
    z0 = p.gw;
    z1 = p.gw + p.nz;
    return true;
  }
  // piece k - full of this block's range of the left-over tiles' tile-planes
  const int t0 = g.full * g.nblk, rest = (g.ntx * g.nty - t0) * p.nz;
  const int lo0 = L * g.per, hi = lo0 + g.per < rest ? lo0 + g.per : rest;
  const int kk = k - g.full;
  const int lo = kk == 0 ? lo0 : (lo0 / p.nz + kk) * p.nz;
  if (lo >= hi) return false;
  const int tr = lo / p.nz, zb = tr * p.nz;
  const int se = hi < zb + p.nz ? hi : zb + p.nz;
  tile = t0 + tr;
  z0 = p.gw + lo - zb;
  z1 = p.gw + se - zb;
  return true;
}

// The block's program, written once for the kernel and for the host emulation: the segments
// of range L (see MarchGeom), one after the other; each thread's CFL maximum spans them all.
template <int RS, int ST, class Exec>
PPK_HD void march_block(Exec &X, const double *__restrict__ Uin, double *__restrict__ Uout,
                        const DevParams &p, const MarchGeom &g, int L, double dt, bool fold,
                        double *lds) {
  using T = MarchThread<RS, ST>;
  X.each([&](T &t) { t.cfl = 0.0; });
  // one call site of the segment program, so that it is inlined once; segment k is computed
  // from (L, k) alone, so that nothing but k is carried across a segment
  int tile, z0, z1;
  for (int k = 0; march_piece(p, g, L, k, tile, z0, z1); ++k)
    march_segment<RS, ST>(X, Uin, Uout, p, g, tile, z0, z1, dt, fold, lds);
}

}  // namespace marchk
}  // namespace PPK_NS
}  // namespace ppk
