"""Write the point-wise solver fixtures: input -> output tables of the oracle's MHD Riemann
solvers on random states, committed so that a change of the oracle's HLLD or 2D HLLD EMF
solver shows as a fixture mismatch.

    python tests/golden/make_golden_solvers.py

hlld_table.npz: 1000 pairs of primitive face states (rho, p, u, v, w, Bx, By, Bz), random
magnitudes of velocity and field per pair, gamma 5/3 and 1.4 alternating; the HLLD flux of each
pair (riemann_hlld, src/shared/RiemannSolvers_MHD.h:133-367).
emf_table.npz: 1000 quadruples of corner states (LL, RL, LR, RR) with their E = u By - v Bx,
velocity and field scales from 0.3 to 15; the 2D HLLD EMF (mag_riemann2d_hlld,
src/shared/RiemannSolvers_MHD.h:398-630), smallc 1e-10.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import textbook as T  # noqa: E402
from oracle import pyoracle as O  # noqa: E402
from ppkmhd_amd import _abi as A  # noqa: E402
from ppkmhd_amd import make_params  # noqa: E402

N = 1000


def params(gamma):
    return make_params(3, 8, 8, 8, gamma0=gamma, riemann=A.RIEMANN_HLLD, smallc=1e-10,
                       float_round=False)


def hlld_table(seed=2024):
    rng = np.random.default_rng(seed)
    gam = np.where(np.arange(N) % 2 == 0, 5.0 / 3.0, 1.4)
    qL = np.zeros((N, 8))
    qR = np.zeros((N, 8))
    F = np.zeros((N, 8))
    for n in range(N):
        bs, vs = rng.uniform(0.05, 2.0), rng.uniform(0.1, 2.0)
        qL[n] = T.random_state(rng, bscale=bs, vscale=vs)
        qR[n] = T.random_state(rng, bscale=bs, vscale=vs)
        F[n] = O.riemann(params(gam[n]), qL[n], qR[n])
    return dict(gamma=gam, qL=qL, qR=qR, flux=F)


def emf_table(seed=4048):
    rng = np.random.default_rng(seed)
    p = params(5.0 / 3.0)
    Q = np.zeros((N, 4, 8))
    E = np.zeros((N, 4))
    out = np.zeros(N)
    for n in range(N):
        vs = [0.3, 1.0, 4.0, 15.0][n % 4]
        for c in range(4):
            Q[n, c] = T.random_state(rng, bscale=vs, vscale=vs)
            E[n, c] = Q[n, c, T.IU] * Q[n, c, T.IB] - Q[n, c, T.IV] * Q[n, c, T.IA]
        out[n] = O.mag_riemann2d_hlld(p, Q[n], E[n])
    return dict(gamma=np.array(5.0 / 3.0), smallc=np.array(1e-10), q=Q, e=E, emf=out)


def main():
    np.savez_compressed(os.path.join(HERE, "hlld_table.npz"), **hlld_table())
    np.savez_compressed(os.path.join(HERE, "emf_table.npz"), **emf_table())
    print("wrote hlld_table.npz, emf_table.npz")


if __name__ == "__main__":
    main()
