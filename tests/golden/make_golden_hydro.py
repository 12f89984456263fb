"""Write the hydro golden fixtures tests/golden/hydro_*.npz from the hydro oracle (CPU
restatement of the reference's implementation-0 Euler step).  The reference ships no golden
vectors for the hydro solvers, so these pin the oracle against regressions and give the GPU
tests fixed inputs/outputs.

    python tests/golden/make_golden_hydro.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import pyhydro as H  # noqa: E402
from ppkmhd_amd import _abi as A  # noqa: E402
from ppkmhd_amd.params import default_problem, make_hydro_params  # noqa: E402


def cases():
    # settings/test_implode_2D.ini at 1/32 of its resolution, the default (approx) solver
    p = make_hydro_params(2, 32, 16, xmax=2.0, ymax=1.0, gamma0=1.666, cfl=0.8, slope_type=2.0,
                          riemann=A.RIEMANN_APPROX, bc=A.BC_DIRICHLET, tEnd=1.0)
    yield "hydro_implode2d_32x16", p, default_problem(A.PROBLEM_IMPLODE, p)
    # settings/test_blast_3D.ini at 1/4 of its resolution, HLLC
    p = make_hydro_params(3, 16, 12, 8, xmax=2.0, ymax=1.5, zmax=1.0, gamma0=1.666, cfl=0.8,
                          slope_type=2.0, riemann=A.RIEMANN_HLLC, bc=A.BC_DIRICHLET, tEnd=50.0)
    yield "hydro_blast3d_16x12x8", p, default_problem(A.PROBLEM_BLAST, p)


def params_dict(p):
    d = {}
    for name, _ in A.Params._fields_:
        v = getattr(p, name)
        d[name] = list(v) if name == "bc" else v
    return d


def main():
    H.build()
    for name, p, pr in cases():
        s = H.Solver(p)
        s.init_problem(pr)
        U0 = s.get_state()
        s.step(0, A.DT_CFL)
        U1 = s.get_state()
        t1 = s.time()
        for _ in range(9):
            s.step(0, A.DT_CFL)
        U10 = s.get_state()
        t10 = s.time()
        np.savez_compressed(os.path.join(HERE, name + ".npz"), U0=U0, U1=U1, U10=U10,
                            time1=np.array(t1[:2]), time10=np.array(t10[:2]),
                            params=np.array(repr(params_dict(p))))
        print(name, U0.shape, t10)


if __name__ == "__main__":
    main()
