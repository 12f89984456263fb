// Hancock predictor (trace) of one cell: face states and edge states at t^{n+1/2}.
// 2D: trace_unsplit_mhd_2d, reference src/muscl/MHDBaseFunctor2D.h:774-1109.
// 3D: trace_unsplit_mhd_3d_simpler, reference src/muscl/MHDBaseFunctor3D.h:688-1114.
// Operand order follows the reference (see mhd_device.hpp).
#pragma once

#include "mhd_device.hpp"

namespace ppk {

// Compact half-step data of one 2D cell: cell centre at t^{n+1/2}, halved slopes, face
// fields at t^{n+1/2} and their transverse slopes.  Every face and edge state of the cell is
// an add/sub of these plus the floors.
template <class R>
struct Half2D {
  R r, pp, u, v, w, A, B, C;
  R drx, dpx, dux, dvx, dwx, dBx, dCx;
  R dry, dpy, duy, dvy, dwy, dAy, dCy;
  R AL, AR, BL, BR;
  R dALy, dBLx, dARy, dBRx;
};

// E_z at the lower-left corner of cell (0,0) from the four cells around it (index m = -1,
// first index x): u, v averaged over the cells, A, B over the two faces touching the corner
template <class R>
PPK_HD R elec_z_2d(R umm, R um0, R u0m, R u00, R vmm, R vm0, R v0m, R v00, R A0m, R A00, R Bm0,
                   R B00) {
  R u = R(0.25) * (umm + um0 + u0m + u00);
  R v = R(0.25) * (vmm + vm0 + v0m + v00);
  R A = R(0.5) * (A0m + A00);
  R B = R(0.5) * (Bm0 + B00);
  return u * B - v * A;
}

// Hancock half step of one 2D cell.  qNb[di][dj] = Q(i+di-1, j+dj-1); bf[di][dj][c] = face
// field c of cell (i+di-1, j+dj-1); Ez = corner fields ELL (i,j), ELR (i,j+1), ERL (i+1,j),
// ERR (i+1,j+1).
template <class R>
PPK_HD void half_step_2d(const St<R> (&qNb)[3][3], const R (&bf)[4][4][3], R ELL, R ELR, R ERL,
                         R ERR, R dtdx, R dtdy, const DevParams &p, Half2D<R> &h) {
  const int C = 1;
  const R gamma = R(p.gamma0);
  const R st = R(p.slope_type);
  const St<R> &q = qNb[C][C];

  R r = q[ID], pp = q[IP], u = q[IU], v = q[IV], w = q[IW], A = q[IA], B = q[IB], Cc = q[IC];
  R AL = bf[C][C][0], AR = bf[C + 1][C][0], BL = bf[C][C][1], BR = bf[C][C + 1][1];

  // hydro slopes (slope_type 1 or 2), halved
  R drx = R(0.5) * limited_slope(st, q[ID], qNb[2][1][ID], qNb[0][1][ID]);
  R dpx = R(0.5) * limited_slope(st, q[IP], qNb[2][1][IP], qNb[0][1][IP]);
  R dux = R(0.5) * limited_slope(st, q[IU], qNb[2][1][IU], qNb[0][1][IU]);
  R dvx = R(0.5) * limited_slope(st, q[IV], qNb[2][1][IV], qNb[0][1][IV]);
  R dwx = R(0.5) * limited_slope(st, q[IW], qNb[2][1][IW], qNb[0][1][IW]);
  R dCx = R(0.5) * limited_slope(st, q[IC], qNb[2][1][IC], qNb[0][1][IC]);
  R dBx = R(0.5) * limited_slope(st, q[IB], qNb[2][1][IB], qNb[0][1][IB]);
  R dry = R(0.5) * limited_slope(st, q[ID], qNb[1][2][ID], qNb[1][0][ID]);
  R dpy = R(0.5) * limited_slope(st, q[IP], qNb[1][2][IP], qNb[1][0][IP]);
  R duy = R(0.5) * limited_slope(st, q[IU], qNb[1][2][IU], qNb[1][0][IU]);
  R dvy = R(0.5) * limited_slope(st, q[IV], qNb[1][2][IV], qNb[1][0][IV]);
  R dwy = R(0.5) * limited_slope(st, q[IW], qNb[1][2][IW], qNb[1][0][IW]);
  R dCy = R(0.5) * limited_slope(st, q[IC], qNb[1][2][IC], qNb[1][0][IC]);
  R dAy = R(0.5) * limited_slope(st, q[IA], qNb[1][2][IA], qNb[1][0][IA]);

  // transverse face-field slopes at (i,j), (i+1,j), (i,j+1) (slope_unsplit_mhd_2d)
  h.dALy = R(0.5) * limited_slope(st, bf[C][C][0], bf[C][C + 1][0], bf[C][C - 1][0]);
  h.dBLx = R(0.5) * limited_slope(st, bf[C][C][1], bf[C + 1][C][1], bf[C - 1][C][1]);
  h.dARy = R(0.5) * limited_slope(st, bf[C + 1][C][0], bf[C + 1][C + 1][0], bf[C + 1][C - 1][0]);
  h.dBRx = R(0.5) * limited_slope(st, bf[C][C + 1][1], bf[C + 1][C + 1][1], bf[C - 1][C + 1][1]);

  R dAx = R(0.5) * (AR - AL);
  R dBy = R(0.5) * (BR - BL);

  R sr0 = (-u * drx - dux * r) * dtdx + (-v * dry - dvy * r) * dtdy;
  const R ir = recip(r);
  R su0 = (-u * dux - divr(dpx, r, ir) - divr(B * dBx, r, ir) - divr(Cc * dCx, r, ir)) * dtdx +
          (-v * duy + divr(B * dAy, r, ir)) * dtdy;
  R sv0 = (-u * dvx + divr(A * dBx, r, ir)) * dtdx +
          (-v * dvy - divr(dpy, r, ir) - divr(A * dAy, r, ir) - divr(Cc * dCy, r, ir)) * dtdy;
  R sw0 = (-u * dwx + divr(A * dCx, r, ir)) * dtdx + (-v * dwy + divr(B * dCy, r, ir)) * dtdy;
  R sp0 = (-u * dpx - dux * gamma * pp) * dtdx + (-v * dpy - dvy * gamma * pp) * dtdy;
  R sA0 = (u * dBy + B * duy - v * dAy - A * dvy) * dtdy;
  R sB0 = (-u * dBx - B * dux + v * dAx + A * dvx) * dtdx;
  R sC0 = (w * dAx + A * dwx - u * dCx - Cc * dux) * dtdx + (-v * dCy - Cc * dvy + w * dBy + B * dwy) * dtdy;
  R sAL0 = +(ELR - ELL) * R(0.5) * dtdy;
  R sAR0 = +(ERR - ERL) * R(0.5) * dtdy;
  R sBL0 = -(ERL - ELL) * R(0.5) * dtdx;
  R sBR0 = -(ERR - ELR) * R(0.5) * dtdx;

  h.r = r + sr0; h.u = u + su0; h.v = v + sv0; h.w = w + sw0; h.pp = pp + sp0;
  h.A = A + sA0; h.B = B + sB0; h.C = Cc + sC0;
  h.AL = AL + sAL0; h.AR = AR + sAR0; h.BL = BL + sBL0; h.BR = BR + sBR0;
  h.drx = drx; h.dpx = dpx; h.dux = dux; h.dvx = dvx; h.dwx = dwx; h.dBx = dBx; h.dCx = dCx;
  h.dry = dry; h.dpy = dpy; h.duy = duy; h.dvy = dvy; h.dwy = dwy; h.dAy = dAy; h.dCy = dCy;
}

// 2D floors: density with smallr, pressure with smallp * density
template <class R>
PPK_HD void floor2(St<R> &s, const DevParams &p) {
  s[ID] = dmax(R(p.smallr), s[ID]);
  s[IP] = dmax(R(p.smallp) * s[ID], s[IP]);
}

// 2D face states: dir 0/1, side -1 = qp (right state at the cell's left interface), +1 = qm
template <class R>
PPK_HD St<R> face_state_2d(const Half2D<R> &h, int dir, int side, const DevParams &p) {
  St<R> s;
  if (dir == 0) {
    if (side < 0) {
      s[ID] = h.r - h.drx; s[IU] = h.u - h.dux; s[IV] = h.v - h.dvx; s[IW] = h.w - h.dwx;
      s[IP] = h.pp - h.dpx; s[IA] = h.AL; s[IB] = h.B - h.dBx; s[IC] = h.C - h.dCx;
    } else {
      s[ID] = h.r + h.drx; s[IU] = h.u + h.dux; s[IV] = h.v + h.dvx; s[IW] = h.w + h.dwx;
      s[IP] = h.pp + h.dpx; s[IA] = h.AR; s[IB] = h.B + h.dBx; s[IC] = h.C + h.dCx;
    }
  } else {
    if (side < 0) {
      s[ID] = h.r - h.dry; s[IU] = h.u - h.duy; s[IV] = h.v - h.dvy; s[IW] = h.w - h.dwy;
      s[IP] = h.pp - h.dpy; s[IA] = h.A - h.dAy; s[IB] = h.BL; s[IC] = h.C - h.dCy;
    } else {
      s[ID] = h.r + h.dry; s[IU] = h.u + h.duy; s[IV] = h.v + h.dvy; s[IW] = h.w + h.dwy;
      s[IP] = h.pp + h.dpy; s[IA] = h.A + h.dAy; s[IB] = h.BR; s[IC] = h.C + h.dCy;
    }
  }
  floor2(s, p);
  return s;
}

// 2D edge (corner) states RT, RB, LT, LB: x is the first transverse direction, y the second
template <class R>
PPK_HD St<R> edge_state_2d(const Half2D<R> &h, int c, const DevParams &p) {
  const bool r1 = (c == IRT || c == IRB);
  const bool r2 = (c == IRT || c == ILT);
  auto cmb = [&](R x0, R d1, R d2) { return x0 + ((r1 ? +d1 : -d1) + (r2 ? d2 : -d2)); };
  St<R> s;
  s[ID] = cmb(h.r, h.drx, h.dry); s[IU] = cmb(h.u, h.dux, h.duy); s[IV] = cmb(h.v, h.dvx, h.dvy);
  s[IW] = cmb(h.w, h.dwx, h.dwy); s[IP] = cmb(h.pp, h.dpx, h.dpy);
  s[IA] = r1 ? (h.AR + (r2 ? +h.dARy : -h.dARy)) : (h.AL + (r2 ? +h.dALy : -h.dALy));
  s[IB] = r2 ? (h.BR + (r1 ? +h.dBRx : -h.dBRx)) : (h.BL + (r1 ? +h.dBLx : -h.dBLx));
  s[IC] = cmb(h.C, h.dCx, h.dCy);
  floor2(s, p);
  return s;
}

// outputs of the 2D trace: qm/qp per direction, edge states RT, RB, LT, LB
template <class R>
struct Trace2D {
  St<R> qm[2], qp[2], qe[4];
};

// whole 2D trace of one cell from its 3x3 primitive and 4x4 face-field neighbourhoods
template <class R>
PPK_HD void trace_cell_2d(const St<R> (&qNb)[3][3], const R (&bf)[4][4][3], R dtdx, R dtdy,
                          const DevParams &p, Trace2D<R> &out) {
  R Ez[2][2];
#pragma unroll
  for (int di = 0; di < 2; ++di)
#pragma unroll
    for (int dj = 0; dj < 2; ++dj) {
      const int cx = 1 + di, cy = 1 + dj;
      Ez[di][dj] = elec_z_2d(qNb[cx - 1][cy - 1][IU], qNb[cx - 1][cy][IU], qNb[cx][cy - 1][IU],
                             qNb[cx][cy][IU], qNb[cx - 1][cy - 1][IV], qNb[cx - 1][cy][IV],
                             qNb[cx][cy - 1][IV], qNb[cx][cy][IV], bf[cx][cy - 1][0],
                             bf[cx][cy][0], bf[cx - 1][cy][1], bf[cx][cy][1]);
    }
  Half2D<R> h;
  half_step_2d(qNb, bf, Ez[0][0], Ez[0][1], Ez[1][0], Ez[1][1], dtdx, dtdy, p, h);
  out.qp[0] = face_state_2d(h, 0, -1, p);
  out.qm[0] = face_state_2d(h, 0, +1, p);
  out.qp[1] = face_state_2d(h, 1, -1, p);
  out.qm[1] = face_state_2d(h, 1, +1, p);
#pragma unroll
  for (int c = 0; c < 4; ++c) out.qe[c] = edge_state_2d(h, c, p);
}

// Compact half-step data of one 3D cell, from which every face state and edge state of the
// cell is an add/sub plus floors (the 6 face and 12 edge states of
// trace_unsplit_mhd_3d_simpler).  R-side face values of a cell equal the L-side values of
// its +1 neighbour, so they are kept here only for the self-contained staged path.
template <class R>
struct Half3D {
  R r, pp, u, v, w, A, B, C;                               // cell centre at t^{n+1/2}
  R drx, dry, drz, dpx, dpy, dpz, dux, duy, duz, dvx, dvy, dvz, dwx, dwy, dwz;  // half slopes
  R dBx, dCx, dAy, dCy, dAz, dBz;
  R AL, AR, BL, BR, CL, CR;                                // face fields at t^{n+1/2}
  R dALy, dALz, dBLx, dBLz, dCLx, dCLy, dARy, dARz, dBRx, dBRz, dCRx, dCRy;  // transverse
};

// inputs: q and its 6 neighbours, the 6 face fields of the cell, 12 transverse face slopes
// (dbf, layout of MHDRunFunctors3D.h:726-738), edge electric fields E[x|y|z][2][2]
// (Ex: ELL ELR ERL ERR, Ey: FLL FLR FRL FRR, Ez: GLL GLR GRL GRR)
// ST: the slope type when fixed at compile time (1: minmod), 0: p.slope_type
template <int ST = 0, class R>
PPK_HD void half_step_3d(const St<R> &q, const St<R> &qPX, const St<R> &qMX,
                                             const St<R> &qPY, const St<R> &qMY, const St<R> &qPZ,
                                             const St<R> &qMZ, const R (&bfNb)[6], const R (&dbf)[12],
                                             const R (&Ex)[4], const R (&Ey)[4], const R (&Ez)[4],
                                             R dtdx, R dtdy, R dtdz, const DevParams &p,
                                             Half3D<R> &h) {
  const R st = ST ? R(ST) : R(p.slope_type);
  const R gamma = R(p.gamma0);
  const bool sl = ST ? true : p.slope12;
  auto S = [&](int v, const St<R> &plus, const St<R> &minus) {
    return sl ? R(0.5) * limited_slope(st, q[v], plus[v], minus[v]) : R(0);
  };
  // hydro slopes, halved (slope_unsplit_hydro_3d, then "*= HALF_F" in the trace)
  const R drx = S(ID, qPX, qMX), dpx = S(IP, qPX, qMX), dux = S(IU, qPX, qMX);
  const R dvx = S(IV, qPX, qMX), dwx = S(IW, qPX, qMX), dCx = S(IC, qPX, qMX), dBx = S(IB, qPX, qMX);
  const R dry = S(ID, qPY, qMY), dpy = S(IP, qPY, qMY), duy = S(IU, qPY, qMY);
  const R dvy = S(IV, qPY, qMY), dwy = S(IW, qPY, qMY), dCy = S(IC, qPY, qMY), dAy = S(IA, qPY, qMY);
  const R drz = S(ID, qPZ, qMZ), dpz = S(IP, qPZ, qMZ), duz = S(IU, qPZ, qMZ);
  const R dvz = S(IV, qPZ, qMZ), dwz = S(IW, qPZ, qMZ), dAz = S(IA, qPZ, qMZ), dBz = S(IB, qPZ, qMZ);

  const R ELL = Ex[0], ELR = Ex[1], ERL = Ex[2], ERR = Ex[3];
  const R FLL = Ey[0], FLR = Ey[1], FRL = Ey[2], FRR = Ey[3];
  const R GLL = Ez[0], GLR = Ez[1], GRL = Ez[2], GRR = Ez[3];

  R r = q[ID], pp = q[IP], u = q[IU], v = q[IV], w = q[IW], A = q[IA], B = q[IB], C = q[IC];
  R AL = bfNb[0], AR = bfNb[1], BL = bfNb[2], BR = bfNb[3], CL = bfNb[4], CR = bfNb[5];

  const R dAx = R(0.5) * (AR - AL);
  const R dBy = R(0.5) * (BR - BL);
  const R dCz = R(0.5) * (CR - CL);

  R sr0 = (-u * drx - dux * r) * dtdx + (-v * dry - dvy * r) * dtdy + (-w * drz - dwz * r) * dtdz;
  const R ir = recip(r);
  R su0 = (-u * dux - divr(dpx + B * dBx + C * dCx, r, ir)) * dtdx +
          (-v * duy + divr(B * dAy, r, ir)) * dtdy + (-w * duz + divr(C * dAz, r, ir)) * dtdz;
  R sv0 = (-u * dvx + divr(A * dBx, r, ir)) * dtdx +
          (-v * dvy - divr(dpy + A * dAy + C * dCy, r, ir)) * dtdy +
          (-w * dvz + divr(C * dBz, r, ir)) * dtdz;
  R sw0 = (-u * dwx + divr(A * dCx, r, ir)) * dtdx + (-v * dwy + divr(B * dCy, r, ir)) * dtdy +
          (-w * dwz - divr(dpz + A * dAz + B * dBz, r, ir)) * dtdz;
  R sp0 = (-u * dpx - dux * gamma * pp) * dtdx + (-v * dpy - dvy * gamma * pp) * dtdy +
          (-w * dpz - dwz * gamma * pp) * dtdz;
  R sA0 = (u * dBy + B * duy - v * dAy - A * dvy) * dtdy + (u * dCz + C * duz - w * dAz - A * dwz) * dtdz;
  R sB0 = (v * dAx + A * dvx - u * dBx - B * dux) * dtdx + (v * dCz + C * dvz - w * dBz - B * dwz) * dtdz;
  R sC0 = (w * dAx + A * dwx - u * dCx - C * dux) * dtdx + (w * dBy + B * dwy - v * dCy - C * dvy) * dtdy;

  R sAL0 = +(GLR - GLL) * dtdy * R(0.5) - (FLR - FLL) * dtdz * R(0.5);
  R sAR0 = +(GRR - GRL) * dtdy * R(0.5) - (FRR - FRL) * dtdz * R(0.5);
  R sBL0 = -(GRL - GLL) * dtdx * R(0.5) + (ELR - ELL) * dtdz * R(0.5);
  R sBR0 = -(GRR - GLR) * dtdx * R(0.5) + (ERR - ERL) * dtdz * R(0.5);
  R sCL0 = +(FRL - FLL) * dtdx * R(0.5) - (ERL - ELL) * dtdy * R(0.5);
  R sCR0 = +(FRR - FLR) * dtdx * R(0.5) - (ERR - ELR) * dtdy * R(0.5);

  h.r = r + sr0; h.u = u + su0; h.v = v + sv0; h.w = w + sw0; h.pp = pp + sp0;
  h.A = A + sA0; h.B = B + sB0; h.C = C + sC0;
  h.AL = AL + sAL0; h.AR = AR + sAR0; h.BL = BL + sBL0; h.BR = BR + sBR0;
  h.CL = CL + sCL0; h.CR = CR + sCR0;
  h.drx = drx; h.dry = dry; h.drz = drz; h.dpx = dpx; h.dpy = dpy; h.dpz = dpz;
  h.dux = dux; h.duy = duy; h.duz = duz; h.dvx = dvx; h.dvy = dvy; h.dvz = dvz;
  h.dwx = dwx; h.dwy = dwy; h.dwz = dwz;
  h.dBx = dBx; h.dCx = dCx; h.dAy = dAy; h.dCy = dCy; h.dAz = dAz; h.dBz = dBz;
  h.dALy = R(0.5) * dbf[0]; h.dALz = R(0.5) * dbf[1]; h.dBLx = R(0.5) * dbf[2];
  h.dBLz = R(0.5) * dbf[3]; h.dCLx = R(0.5) * dbf[4]; h.dCLy = R(0.5) * dbf[5];
  h.dARy = R(0.5) * dbf[6]; h.dARz = R(0.5) * dbf[7]; h.dBRx = R(0.5) * dbf[8];
  h.dBRz = R(0.5) * dbf[9]; h.dCRx = R(0.5) * dbf[10]; h.dCRy = R(0.5) * dbf[11];
}

template <class R>
PPK_HD void floor3(St<R> &s, const DevParams &p) {
  s[ID] = dmax(R(p.smallr), s[ID]);
  s[IP] = dmax(R(p.smallp), s[IP]);  // 3D floors with smallp (not smallp*rho)
}

// face states: dir 0/1/2, side -1 = qp (right state at left interface), +1 = qm
template <class R>
PPK_HD St<R> face_state_3d(const Half3D<R> &h, int dir, int side,
                                               const DevParams &p) {
  St<R> s;
  if (dir == 0) {
    if (side < 0) {
      s[ID] = h.r - h.drx; s[IU] = h.u - h.dux; s[IV] = h.v - h.dvx; s[IW] = h.w - h.dwx;
      s[IP] = h.pp - h.dpx; s[IA] = h.AL; s[IB] = h.B - h.dBx; s[IC] = h.C - h.dCx;
    } else {
      s[ID] = h.r + h.drx; s[IU] = h.u + h.dux; s[IV] = h.v + h.dvx; s[IW] = h.w + h.dwx;
      s[IP] = h.pp + h.dpx; s[IA] = h.AR; s[IB] = h.B + h.dBx; s[IC] = h.C + h.dCx;
    }
  } else if (dir == 1) {
    if (side < 0) {
      s[ID] = h.r - h.dry; s[IU] = h.u - h.duy; s[IV] = h.v - h.dvy; s[IW] = h.w - h.dwy;
      s[IP] = h.pp - h.dpy; s[IA] = h.A - h.dAy; s[IB] = h.BL; s[IC] = h.C - h.dCy;
    } else {
      s[ID] = h.r + h.dry; s[IU] = h.u + h.duy; s[IV] = h.v + h.dvy; s[IW] = h.w + h.dwy;
      s[IP] = h.pp + h.dpy; s[IA] = h.A + h.dAy; s[IB] = h.BR; s[IC] = h.C + h.dCy;
    }
  } else {
    if (side < 0) {
      s[ID] = h.r - h.drz; s[IU] = h.u - h.duz; s[IV] = h.v - h.dvz; s[IW] = h.w - h.dwz;
      s[IP] = h.pp - h.dpz; s[IA] = h.A - h.dAz; s[IB] = h.B - h.dBz; s[IC] = h.CL;
    } else {
      s[ID] = h.r + h.drz; s[IU] = h.u + h.duz; s[IV] = h.v + h.dvz; s[IW] = h.w + h.dwz;
      s[IP] = h.pp + h.dpz; s[IA] = h.A + h.dAz; s[IB] = h.B + h.dBz; s[IC] = h.CR;
    }
  }
  floor3(s, p);
  return s;
}

// edge states: edge direction e (0 X, 1 Y, 2 Z), corner c in IRT, IRB, ILT, ILB
template <class R>
PPK_HD St<R> edge_state_3d(const Half3D<R> &h, int e, int c,
                                               const DevParams &p) {
  St<R> s;
  // signs of the two transverse half-slopes for RT, RB, LT, LB: (s1, s2)
  const bool r1 = (c == IRT || c == IRB);  // first transverse: +
  const bool r2 = (c == IRT || c == ILT);  // second transverse: +
  if (e == 0) {  // X edge: transverse y (first) and z (second)
    auto cmb = [&](R x0, R d1, R d2) { return x0 + ((r1 ? +d1 : -d1) + (r2 ? d2 : -d2)); };
    s[ID] = cmb(h.r, h.dry, h.drz); s[IU] = cmb(h.u, h.duy, h.duz); s[IV] = cmb(h.v, h.dvy, h.dvz);
    s[IW] = cmb(h.w, h.dwy, h.dwz); s[IP] = cmb(h.pp, h.dpy, h.dpz); s[IA] = cmb(h.A, h.dAy, h.dAz);
    s[IB] = r1 ? (h.BR + (r2 ? +h.dBRz : -h.dBRz)) : (h.BL + (r2 ? +h.dBLz : -h.dBLz));
    s[IC] = r2 ? (h.CR + (r1 ? +h.dCRy : -h.dCRy)) : (h.CL + (r1 ? +h.dCLy : -h.dCLy));
  } else if (e == 1) {  // Y edge: transverse x (first) and z (second)
    auto cmb = [&](R x0, R d1, R d2) { return x0 + ((r1 ? +d1 : -d1) + (r2 ? d2 : -d2)); };
    s[ID] = cmb(h.r, h.drx, h.drz); s[IU] = cmb(h.u, h.dux, h.duz); s[IV] = cmb(h.v, h.dvx, h.dvz);
    s[IW] = cmb(h.w, h.dwx, h.dwz); s[IP] = cmb(h.pp, h.dpx, h.dpz);
    s[IA] = r1 ? (h.AR + (r2 ? +h.dARz : -h.dARz)) : (h.AL + (r2 ? +h.dALz : -h.dALz));
    s[IB] = cmb(h.B, h.dBx, h.dBz);
    s[IC] = r2 ? (h.CR + (r1 ? +h.dCRx : -h.dCRx)) : (h.CL + (r1 ? +h.dCLx : -h.dCLx));
  } else {  // Z edge: transverse x (first) and y (second)
    auto cmb = [&](R x0, R d1, R d2) { return x0 + ((r1 ? +d1 : -d1) + (r2 ? d2 : -d2)); };
    s[ID] = cmb(h.r, h.drx, h.dry); s[IU] = cmb(h.u, h.dux, h.duy); s[IV] = cmb(h.v, h.dvx, h.dvy);
    s[IW] = cmb(h.w, h.dwx, h.dwy); s[IP] = cmb(h.pp, h.dpx, h.dpy);
    s[IA] = r1 ? (h.AR + (r2 ? +h.dARy : -h.dARy)) : (h.AL + (r2 ? +h.dALy : -h.dALy));
    s[IB] = r2 ? (h.BR + (r1 ? +h.dBRx : -h.dBRx)) : (h.BL + (r1 ? +h.dBLx : -h.dBLx));
    s[IC] = cmb(h.C, h.dCx, h.dCy);
  }
  floor3(s, p);
  return s;
}

}  // namespace ppk
