"""Pin the oracle's point-wise pieces against the mathematics (CPU only).

Oracle functions and the checks that pin them:
  riemann_hlld         textbook HLLD (Miyoshi & Kusano 2005) via jump conditions, consistency,
                       upwinding, mirror symmetry, the B = 0 contact
  riemann_hll / llf    textbook HLL / LLF formulas, consistency
  mag_riemann2d_hlld   branch form of the same solver (independent restatement), uniform
                       states, supersonic corners
  compute_emf          v x B for uniform edge states, for each of the three EMF directions
  find_speed_fast      textbook fast speed
  slope limiter        textbook minmod and MC
"""
import numpy as np
import pytest

import textbook as T
from ppkmhd_amd import _abi as A
from ppkmhd_amd import make_params


def P(gamma=5 / 3, riemann=A.RIEMANN_HLLD, slope_type=1.0, smallc=1e-10):
    return make_params(3, 8, 8, 8, gamma0=gamma, riemann=riemann, slope_type=slope_type,
                       smallc=smallc, float_round=False)


def rel(a, b, floor=1.0):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(floor, float(np.max(np.abs(b)))))


@pytest.mark.parametrize("seed", range(6))
def test_hlld_matches_textbook(oracle, seed):
    rng = np.random.default_rng(seed)
    p = P(gamma=[5 / 3, 1.4][seed % 2])
    worst = 0.0
    for _ in range(300):
        qL = T.random_state(rng, bscale=rng.uniform(0.05, 2.0), vscale=rng.uniform(0.1, 2.0))
        qR = T.random_state(rng, bscale=rng.uniform(0.05, 2.0), vscale=rng.uniform(0.1, 2.0))
        f = oracle.riemann(p, qL, qR)
        g = T.hlld(qL, qR, p.gamma0)
        worst = max(worst, rel(f, g))
    assert worst < 1e-11, worst


def test_hlld_consistency(oracle):
    """F(q, q) is the physical flux of q (closed form)."""
    rng = np.random.default_rng(11)
    p = P()
    for _ in range(200):
        q = T.random_state(rng)
        f = oracle.riemann(p, q, q)
        _, F = T.physical_flux(q, p.gamma0)
        assert rel(f, F) < 1e-13


def test_hlld_supersonic_upwinding(oracle):
    """All waves right-moving -> F(q_L); all left-moving -> F(q_R)."""
    rng = np.random.default_rng(12)
    p = P()
    for _ in range(100):
        qL, qR = T.random_state(rng), T.random_state(rng)
        qR[T.IA] = qL[T.IA]
        for sgn in (1, -1):
            a, b = qL.copy(), qR.copy()
            a[T.IU] = b[T.IU] = sgn * 40.0
            f = oracle.riemann(p, a, b)
            _, F = T.physical_flux(a if sgn > 0 else b, p.gamma0)
            assert rel(f, F) < 1e-13


def test_hlld_mirror_symmetry(oracle):
    """x -> -x: swap L/R, negate u and Bx; odd flux components change sign."""
    rng = np.random.default_rng(13)
    p = P()
    M = np.array([1, 1, -1, 1, 1, -1, 1, 1.0])
    S = np.array([-1, -1, 1, -1, -1, 1, -1, -1.0])
    for _ in range(200):
        qL, qR = T.random_state(rng), T.random_state(rng)
        f = oracle.riemann(p, qL, qR)
        g = oracle.riemann(p, M * qR, M * qL)
        assert rel(g, S * f) < 1e-12


def test_hlld_b0_resolves_contact(oracle):
    """B = 0, equal pressure and normal velocity u0: the exact solution is a contact (with a
    shear of v, w) moving at u0, and HLLD, like HLLC, returns the upwind side's physical flux;
    at u0 = 0 the mass flux is zero.  HLL smears the same contact (the check can tell)."""
    rng = np.random.default_rng(14)
    p, phll = P(), P(riemann=A.RIEMANN_HLL)
    hll_worst = 0.0
    for n in range(200):
        u0 = 0.0 if n < 50 else rng.uniform(-0.5, 0.5)
        pr = rng.uniform(0.5, 3.0)
        qL, qR = T.random_state(rng, bscale=0.0), T.random_state(rng, bscale=0.0)
        qL[T.IP] = qR[T.IP] = pr
        qL[T.IU] = qR[T.IU] = u0
        f = oracle.riemann(p, qL, qR)
        _, F = T.physical_flux(qL if u0 >= 0 else qR, p.gamma0)
        assert rel(f, F) < 1e-13, (n, f, F)
        if u0 == 0.0:
            assert f[T.ID] == 0.0
        hll_worst = max(hll_worst, rel(oracle.riemann(phll, qL, qR), F))
    assert hll_worst > 1e-2


@pytest.mark.parametrize("riemann", [A.RIEMANN_HLL, A.RIEMANN_LLF])
def test_hll_llf_textbook(oracle, riemann):
    rng = np.random.default_rng(14)
    p = P(riemann=riemann)
    for _ in range(200):
        qL, qR = T.random_state(rng), T.random_state(rng)
        bx = 0.5 * (qL[T.IA] + qR[T.IA])
        a, b = qL.copy(), qR.copy()
        a[T.IA] = b[T.IA] = bx
        UL, FL = T.physical_flux(a, p.gamma0)
        UR, FR = T.physical_flux(b, p.gamma0)
        if riemann == A.RIEMANN_HLL:
            c = max(T.fast_speed(a, p.gamma0), T.fast_speed(b, p.gamma0))
            sl = min(min(a[T.IU], b[T.IU]) - c, 0.0)
            sr = max(max(a[T.IU], b[T.IU]) + c, 0.0)
            g = (sr * FL - sl * FR + sr * sl * (UR - UL)) / (sr - sl)
        else:
            c = max(T.fast_speed(a, p.gamma0) + abs(a[T.IU]), T.fast_speed(b, p.gamma0) + abs(b[T.IU]))
            g = 0.5 * (FL + FR) - 0.5 * c * (UR - UL)
        f = oracle.riemann(p, qL, qR)
        assert rel(f, g) < 1e-12
        f2 = oracle.riemann(p, qL, qL)
        assert rel(f2, T.physical_flux(qL, p.gamma0)[1]) < 1e-13


def test_fast_speed_textbook(oracle):
    rng = np.random.default_rng(15)
    p = P()
    for _ in range(200):
        q = T.random_state(rng)
        for d in range(3):
            assert abs(oracle.find_speed_fast(p, q, d) - T.fast_speed(q, p.gamma0, d)) < 1e-12 * (
                1 + T.fast_speed(q, p.gamma0, d))


@pytest.mark.parametrize("slope_type", [1.0, 2.0])
def test_limiter_textbook(oracle, slope_type):
    rng = np.random.default_rng(16)
    p = P(slope_type=slope_type)
    for _ in range(500):
        q, qp, qm = rng.normal(size=3)
        if rng.uniform() < 0.1:
            qp = q  # flat side
        assert oracle.slope(p, q, qp, qm) == pytest.approx(float(T.limiter(q, qp, qm, slope_type)),
                                                           abs=1e-15)


def _corner_states(rng, scale=1.0):
    qs = []
    for _ in range(4):
        q = T.random_state(rng, bscale=scale, vscale=scale)
        qs.append(q)
    return qs


@pytest.mark.parametrize("seed", range(4))
def test_mag_riemann2d_blend_equals_branch_form(oracle, seed):
    rng = np.random.default_rng(100 + seed)
    p = P(smallc=1e-10)
    worst = 0.0
    for it in range(300):
        vs = [0.3, 1.0, 4.0, 15.0][it % 4]
        qLL, qRL, qLR, qRR = _corner_states(rng, vs)
        E = [q[T.IU] * q[T.IB] - q[T.IV] * q[T.IA] for q in (qLL, qRL, qLR, qRR)]
        got = oracle.mag_riemann2d_hlld(p, np.array([qLL, qRL, qLR, qRR]), np.array(E))
        ref = T.mag_riemann2d_branchy(qLL, qRL, qLR, qRR, *E, p.gamma0, p.smallc)
        worst = max(worst, abs(got - ref) / (1.0 + abs(ref)))
    assert worst < 1e-11, worst


@pytest.mark.parametrize("edir", [0, 1, 2])
def test_compute_emf_uniform_is_v_cross_b(oracle, edir):
    """Four identical edge states: the EMF is the component of v x B along the edge
    (the reference's E = v x B sign convention)."""
    rng = np.random.default_rng(200 + edir)
    p = P()
    for _ in range(100):
        q = T.random_state(rng)
        e = oracle.compute_emf(p, edir, np.array([q, q, q, q]))
        v = q[T.IU:T.IU + 3]
        B = q[T.IA:T.IA + 3]
        vxb = np.cross(v, B)
        assert e == pytest.approx(vxb[edir], rel=1e-11, abs=1e-12)


def test_mag_riemann2d_supersonic_corner(oracle):
    """All four states moving fast towards +x, +y: the EMF is the upwind ELL."""
    rng = np.random.default_rng(300)
    p = P()
    for _ in range(50):
        qs = _corner_states(rng)
        for q in qs:
            q[T.IU] += 50.0
            q[T.IV] += 50.0
        E = [q[T.IU] * q[T.IB] - q[T.IV] * q[T.IA] for q in qs]
        got = oracle.mag_riemann2d_hlld(p, np.array(qs), np.array(E))
        assert got == E[0]
