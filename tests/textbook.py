"""Independent textbook formulas used to pin the oracle (test helper, numpy).

These are written from the published mathematics, not from the reference's code, and take a
different computational route where one exists (e.g. HLLD fluxes via F*_a = F_a + S_a(U*_a -
U_a) instead of the physical flux of the sampled state), so that a transcription slip in the
oracle (a dropped term, a wrong sign or index, a swapped operand) shows up as a mismatch.

State vectors use the reference's variable order: rho, p, u, v, w, Bx, By, Bz (primitive) and
rho, E, rho*u, rho*v, rho*w, Bx, By, Bz (conservative).
"""
import numpy as np

ID, IP, IU, IV, IW, IA, IB, IC = range(8)


def physical_flux(q, gamma):
    """Ideal MHD x-flux and conservative vector of a primitive state (textbook)."""
    r, p, u, v, w, bx, by, bz = q
    b2 = bx * bx + by * by + bz * bz
    e = p / (gamma - 1.0) + 0.5 * r * (u * u + v * v + w * w) + 0.5 * b2
    pt = p + 0.5 * b2
    vb = u * bx + v * by + w * bz
    U = np.array([r, e, r * u, r * v, r * w, bx, by, bz])
    F = np.array([r * u, (e + pt) * u - bx * vb, r * u * u + pt - bx * bx, r * u * v - bx * by,
                  r * u * w - bx * bz, 0.0, by * u - bx * v, bz * u - bx * w])
    return U, F


def fast_speed(q, gamma, d=0):
    """Fast magnetosonic speed along axis d from the quadratic cf^2 =
    (a^2 + b^2 + sqrt((a^2 + b^2)^2 - 4 a^2 b_d^2)) / 2, a^2 = gamma p / rho, b = B/sqrt(rho)."""
    r, p = q[ID], q[IP]
    B = np.array(q[IA:IA + 3])
    a2 = gamma * p / r
    b2 = (B @ B) / r
    bd2 = B[d] ** 2 / r
    return np.sqrt(0.5 * (a2 + b2 + np.sqrt((a2 + b2) ** 2 - 4.0 * a2 * bd2)))


def minmod(a, b):
    return np.where(a * b > 0, np.sign(a) * np.minimum(np.abs(a), np.abs(b)), 0.0)


def limiter(q, qp, qm, slope_type):
    """Textbook minmod (slope_type 1) and monotonized-central (slope_type 2) slopes."""
    dl, dr, dc = q - qm, qp - q, 0.5 * (qp - qm)
    if slope_type == 1:
        return minmod(dl, dr)
    mc = np.sign(dc) * np.minimum(np.minimum(2 * np.abs(dl), 2 * np.abs(dr)), np.abs(dc))
    return np.where(dl * dr > 0, mc, 0.0)


def hlld(qL, qR, gamma):
    """HLLD flux, Miyoshi & Kusano, J. Comput. Phys. 208 (2005) 315, eqs. (38)-(66), with the
    outer wave-speed estimate S_L = min(u_L, u_R) - max(cf_L, cf_R) (and S_R symmetric) that the
    reference uses.  Fluxes are assembled with the jump conditions F*_a = F_a + S_a (U*_a - U_a),
    F**_a = F*_a + S*_a (U**_a - U*_a)."""
    qL = np.array(qL, dtype=float)
    qR = np.array(qR, dtype=float)
    bx = 0.5 * (qL[IA] + qR[IA])
    qL[IA] = bx
    qR[IA] = bx
    UL, FL = physical_flux(qL, gamma)
    UR, FR = physical_flux(qR, gamma)
    cf = max(fast_speed(qL, gamma), fast_speed(qR, gamma))
    SL = min(qL[IU], qR[IU]) - cf
    SR = max(qL[IU], qR[IU]) + cf
    if SL > 0:
        return FL
    if SR < 0:
        return FR
    rL, uL, rR, uR = qL[ID], qL[IU], qR[ID], qR[IU]
    ptL = qL[IP] + 0.5 * (qL[IA:IA + 3] @ qL[IA:IA + 3])
    ptR = qR[IP] + 0.5 * (qR[IA:IA + 3] @ qR[IA:IA + 3])
    den = (SR - uR) * rR - (SL - uL) * rL
    SM = ((SR - uR) * rR * uR - (SL - uL) * rL * uL - ptR + ptL) / den          # (38)
    pts = ((SR - uR) * rR * ptL - (SL - uL) * rL * ptR +
           rL * rR * (SR - uR) * (SL - uL) * (uR - uL)) / den                   # (41)

    def star(q, U, S):
        r, u = q[ID], q[IU]
        v, w, by, bz = q[IV], q[IW], q[IB], q[IC]
        rs = r * (S - u) / (S - SM)                                              # (43)
        dd = r * (S - u) * (S - SM) - bx * bx
        if bx * bx > 0 and abs(dd / (bx * bx)) <= 1e-8:   # degenerate case, M&K section 4.1
            vs, ws, bys, bzs = v, w, by, bz
        else:
            vs = v - bx * by * (SM - u) / dd                                      # (44)
            ws = w - bx * bz * (SM - u) / dd                                      # (46)
            bys = by * (r * (S - u) ** 2 - bx * bx) / dd                          # (45)
            bzs = bz * (r * (S - u) ** 2 - bx * bx) / dd                          # (47)
        vb = u * bx + v * by + w * bz
        vbs = SM * bx + vs * bys + ws * bzs
        es = ((S - u) * U[IP] - (q[IP] + 0.5 * (q[IA:IA + 3] @ q[IA:IA + 3])) * u + pts * SM +
              bx * (vb - vbs)) / (S - SM)                                        # (48)
        Us = np.array([rs, es, rs * SM, rs * vs, rs * ws, bx, bys, bzs])
        return Us, rs, vs, ws, bys, bzs, vbs

    ULs, rLs, vLs, wLs, byLs, bzLs, vbLs = star(qL, UL, SL)
    URs, rRs, vRs, wRs, byRs, bzRs, vbRs = star(qR, UR, SR)
    FLs = FL + SL * (ULs - UL)
    FRs = FR + SR * (URs - UR)
    sL, sR = np.sqrt(rLs), np.sqrt(rRs)
    SLs = SM - abs(bx) / sL                                                      # (51)
    SRs = SM + abs(bx) / sR
    if SLs > 0:
        return FLs
    if SRs < 0:
        return FRs
    sg = 1.0 if bx >= 0 else -1.0
    vss = (sL * vLs + sR * vRs + (byRs - byLs) * sg) / (sL + sR)                 # (59)
    wss = (sL * wLs + sR * wRs + (bzRs - bzLs) * sg) / (sL + sR)                 # (60)
    byss = (sL * byRs + sR * byLs + sL * sR * (vRs - vLs) * sg) / (sL + sR)      # (61)
    bzss = (sL * bzRs + sR * bzLs + sL * sR * (wRs - wLs) * sg) / (sL + sR)      # (62)
    vbss = SM * bx + vss * byss + wss * bzss
    eLss = ULs[IP] - sL * (vbLs - vbss) * sg                                     # (63)
    eRss = URs[IP] + sR * (vbRs - vbss) * sg
    ULss = np.array([rLs, eLss, rLs * SM, rLs * vss, rLs * wss, bx, byss, bzss])
    URss = np.array([rRs, eRss, rRs * SM, rRs * vss, rRs * wss, bx, byss, bzss])
    if SM > 0:
        return FLs + SLs * (ULss - ULs)
    return FRs + SRs * (URss - URs)


def mag_riemann2d_branchy(qLL, qRL, qLR, qRR, ELL, ERL, ELR, ERR, gamma, smallc):
    """2D HLLD EMF (Londrillo & Del Zanna upwind constrained transport as in DUMSES
    cmp_mag_flx) in its branch form: region selection with if/else on SB, ST, SL, SR instead of
    the 0/1 blend."""
    def sp(q, d):
        return fast_speed(q, gamma, d)
    qs = (qLL, qLR, qRL, qRR)
    u = [q[IU] for q in qs]
    v = [q[IV] for q in qs]
    cx = max(sp(q, 0) for q in qs)
    cy = max(sp(q, 1) for q in qs)
    SL, SR, SB, ST = min(u) - cx, max(u) + cx, min(v) - cy, max(v) + cy
    rLL, uLL, vLL, aLL, bLL = qLL[ID], qLL[IU], qLL[IV], qLL[IA], qLL[IB]
    rLR, uLR, vLR, aLR, bLR = qLR[ID], qLR[IU], qLR[IV], qLR[IA], qLR[IB]
    rRL, uRL, vRL, aRL, bRL = qRL[ID], qRL[IU], qRL[IV], qRL[IA], qRL[IB]
    rRR, uRR, vRR, aRR, bRR = qRR[ID], qRR[IU], qRR[IV], qRR[IA], qRR[IB]
    pt = [q[IP] + 0.5 * (q[IA:IA + 3] @ q[IA:IA + 3]) for q in (qLL, qLR, qRL, qRR)]
    PLL, PLR, PRL, PRR = pt
    rcLLx, rcRLx, rcLRx, rcRRx = rLL * (uLL - SL), rRL * (SR - uRL), rLR * (uLR - SL), rRR * (SR - uRR)
    rcLLy, rcLRy, rcRLy, rcRRy = rLL * (vLL - SB), rLR * (ST - vLR), rRL * (vRL - SB), rRR * (ST - vRR)
    us = (rcLLx * uLL + rcLRx * uLR + rcRLx * uRL + rcRRx * uRR + (PLL - PRL + PLR - PRR)) / \
        (rcLLx + rcLRx + rcRLx + rcRRx)
    vs = (rcLLy * vLL + rcLRy * vLR + rcRLy * vRL + rcRRy * vRR + (PLL - PLR + PRL - PRR)) / \
        (rcLLy + rcLRy + rcRLy + rcRRy)

    def st(r, uu, vv, a, b, Sx, Sy):
        fx, fy = (Sx - uu) / (Sx - us), (Sy - vv) / (Sy - vs)
        return r * fx, b * fx, r * fy, a * fy, r * fx * fy
    rLLx, BLL, rLLy, ALL, rLLs = st(rLL, uLL, vLL, aLL, bLL, SL, SB)
    rLRx, BLR, rLRy, ALR, rLRs = st(rLR, uLR, vLR, aLR, bLR, SL, ST)
    rRLx, BRL, rRLy, ARL, rRLs = st(rRL, uRL, vRL, aRL, bRL, SR, SB)
    rRRx, BRR, rRRy, ARR, rRRs = st(rRR, uRR, vRR, aRR, bRR, SR, ST)
    EsLLx, EsLLy, EsLL = us * BLL - vLL * aLL, uLL * bLL - vs * ALL, us * BLL - vs * ALL
    EsLRx, EsLRy, EsLR = us * BLR - vLR * aLR, uLR * bLR - vs * ALR, us * BLR - vs * ALR
    EsRLx, EsRLy, EsRL = us * BRL - vRL * aRL, uRL * bRL - vs * ARL, us * BRL - vs * ARL
    EsRRx, EsRRy, EsRR = us * BRR - vRR * aRR, uRR * bRR - vs * ARR, us * BRR - vs * ARR
    cL = max(abs(aLR) / np.sqrt(rLRx), abs(ALR) / np.sqrt(rLRs), abs(aLL) / np.sqrt(rLLx),
             abs(ALL) / np.sqrt(rLLs), smallc)
    cR = max(abs(aRR) / np.sqrt(rRRx), abs(ARR) / np.sqrt(rRRs), abs(aRL) / np.sqrt(rRLx),
             abs(ARL) / np.sqrt(rRLs), smallc)
    cB = max(abs(bLL) / np.sqrt(rLLy), abs(BLL) / np.sqrt(rLLs), abs(bRL) / np.sqrt(rRLy),
             abs(BRL) / np.sqrt(rRLs), smallc)
    cT = max(abs(bLR) / np.sqrt(rLRy), abs(BLR) / np.sqrt(rLRs), abs(bRR) / np.sqrt(rRRy),
             abs(BRR) / np.sqrt(rRRs), smallc)
    SAL, SAR = min(us - cL, 0.0), max(us + cR, 0.0)
    SAB, SAT = min(vs - cB, 0.0), max(vs + cT, 0.0)
    AT = (SAR * ARR - SAL * ALR) / (SAR - SAL)
    AB = (SAR * ARL - SAL * ALL) / (SAR - SAL)
    BR = (SAT * BRR - SAB * BRL) / (SAT - SAB)
    BL = (SAT * BLR - SAB * BLL) / (SAT - SAB)
    if SB > 0:
        if SL > 0:
            return ELL
        if SR < 0:
            return ERL
        return (SAR * EsLLx - SAL * EsRLx + SAR * SAL * (bRL - bLL)) / (SAR - SAL)
    if ST < 0:
        if SL > 0:
            return ELR
        if SR < 0:
            return ERR
        return (SAR * EsLRx - SAL * EsRRx + SAR * SAL * (bRR - bLR)) / (SAR - SAL)
    if SL > 0:
        return (SAT * EsLLy - SAB * EsLRy - SAT * SAB * (aLR - aLL)) / (SAT - SAB)
    if SR < 0:
        return (SAT * EsRLy - SAB * EsRRy - SAT * SAB * (aRR - aRL)) / (SAT - SAB)
    return ((SAL * SAB * EsRR - SAL * SAT * EsRL - SAR * SAB * EsLR + SAR * SAT * EsLL)
            / (SAR - SAL) / (SAT - SAB) - SAT * SAB / (SAT - SAB) * (AT - AB)
            + SAR * SAL / (SAR - SAL) * (BR - BL))


def random_state(rng, bscale=1.0, vscale=1.0):
    """Random physical primitive state (positive density and pressure)."""
    return np.array([rng.uniform(0.2, 3.0), rng.uniform(0.05, 5.0), *rng.normal(0, vscale, 3),
                     *rng.normal(0, bscale, 3)])


def fill_ghosts_textbook(U, nx, ny, nz, gw, bc, dim, mhd=True):
    """Ghost fill defined directly: faces in the order x, y, z, each over the whole transverse
    extent; periodic wraps, Neumann copies the first/last interior cell, Dirichlet mirrors
    with the normal velocity (and, for MHD, the normal field) negated."""
    U = U.copy()
    ns = (nx, ny, nz)
    axes = {0: 3, 1: 2, 2: 1}  # numpy axis of x, y, z in (var, k, j, i)
    for d in range(dim):
        ax = axes[d]
        n = ns[d]
        for upper in (0, 1):
            kind = bc[2 * d + upper]
            for g in range(gw):
                c = g + n + gw if upper else g
                if kind == 3:
                    c0 = c - n if upper else c + n
                elif kind == 2:
                    c0 = n + gw - 1 if upper else gw
                else:
                    c0 = 2 * n + 2 * gw - 1 - c if upper else 2 * gw - 1 - c
                src = np.take(U, c0, axis=ax).copy()
                if kind == 1:
                    src[IU + d] *= -1
                    if mhd:
                        src[IA + d] *= -1
                idx = [slice(None)] * 4
                idx[ax] = c
                U[tuple(idx)] = src
    return U


# ---- Euler equations ---------------------------------------------------------------------------
def euler_flux(q, gamma):
    """Physical x-flux of the Euler equations for a primitive state in the reference's variable
    order (rho, p, u, v[, w]), returned in the same order (mass, energy, x, y[, z] momentum)."""
    q = np.asarray(q, dtype=np.float64)
    r, p, u = q[0], q[1], q[2]
    vt = q[3:]
    E = p / (gamma - 1) + 0.5 * r * (u * u + np.sum(vt * vt))
    return np.array([r * u, u * (E + p), r * u * u + p, *(r * u * vt)])


def exact_riemann(rl, ul, pl, rr, ur, pr, gamma, xi):
    """Exact solution (rho, u, p) of the 1D Euler Riemann problem sampled at x/t = xi
    (Toro, Riemann Solvers and Numerical Methods for Fluid Dynamics, ch. 4: pressure function,
    Newton iteration for p*, sampling of the wave pattern)."""
    g = gamma
    cl, cr = np.sqrt(g * pl / rl), np.sqrt(g * pr / rr)

    def fk(p, rk, pk, ck):
        if p > pk:  # shock
            A, B = 2 / ((g + 1) * rk), (g - 1) / (g + 1) * pk
            f = (p - pk) * np.sqrt(A / (p + B))
            df = np.sqrt(A / (p + B)) * (1 - (p - pk) / (2 * (B + p)))
        else:  # rarefaction
            f = 2 * ck / (g - 1) * ((p / pk) ** ((g - 1) / (2 * g)) - 1)
            df = 1 / (rk * ck) * (p / pk) ** (-(g + 1) / (2 * g))
        return f, df

    p = max(1e-8, 0.5 * (pl + pr))
    for _ in range(100):
        f1, d1 = fk(p, rl, pl, cl)
        f2, d2 = fk(p, rr, pr, cr)
        dp = (f1 + f2 + ur - ul) / (d1 + d2)
        p_new = max(1e-12, p - dp)
        if abs(p_new - p) < 1e-14 * (p_new + p):
            p = p_new
            break
        p = p_new
    f1, _ = fk(p, rl, pl, cl)
    f2, _ = fk(p, rr, pr, cr)
    us = 0.5 * (ul + ur) + 0.5 * (f2 - f1)
    if xi <= us:  # left of the contact
        if p > pl:
            sl = ul - cl * np.sqrt((g + 1) / (2 * g) * p / pl + (g - 1) / (2 * g))
            if xi <= sl:
                return rl, ul, pl
            return rl * (p / pl + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * p / pl + 1), us, p
        csl = cl * (p / pl) ** ((g - 1) / (2 * g))
        if xi <= ul - cl:
            return rl, ul, pl
        if xi >= us - csl:
            return rl * (p / pl) ** (1 / g), us, p
        c = 2 / (g + 1) * (cl + (g - 1) / 2 * (ul - xi))
        u = 2 / (g + 1) * (cl + (g - 1) / 2 * ul + xi)
        r = rl * (c / cl) ** (2 / (g - 1))
        return r, u, pl * (c / cl) ** (2 * g / (g - 1))
    if p > pr:
        sr = ur + cr * np.sqrt((g + 1) / (2 * g) * p / pr + (g - 1) / (2 * g))
        if xi >= sr:
            return rr, ur, pr
        return rr * (p / pr + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * p / pr + 1), us, p
    csr = cr * (p / pr) ** ((g - 1) / (2 * g))
    if xi >= ur + cr:
        return rr, ur, pr
    if xi <= us + csr:
        return rr * (p / pr) ** (1 / g), us, p
    c = 2 / (g + 1) * (cr - (g - 1) / 2 * (ur - xi))
    u = 2 / (g + 1) * (-cr + (g - 1) / 2 * ur + xi)
    r = rr * (c / cr) ** (2 / (g - 1))
    return r, u, pr * (c / cr) ** (2 * g / (g - 1))
