"""Write the golden fixtures tests/golden/*.npz from the oracle (CPU restatement of the
reference's implementation-0 step).  The reference ships no golden vectors for this path, so
these pin the oracle against regressions and give the GPU tests fixed inputs/outputs.

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import pyoracle as O  # noqa: E402
from ppkmhd_amd import _abi as A  # noqa: E402
from ppkmhd_amd import default_problem, make_params  # noqa: E402


def cases():
    p = make_params(2, 16, 16, gamma0=5 / 3, cfl=0.8, slope_type=1.0, tEnd=0.5)
    yield "ot2d_16", p, default_problem(A.PROBLEM_ORSZAG_TANG, p)
    p = make_params(3, 8, 8, 8, gamma0=5 / 3, cfl=0.8, slope_type=1.0, tEnd=1.0)
    pr = default_problem(A.PROBLEM_BLAST, p)
    pr.blast_radius = 0.2
    pr.blast_density_out = 1.0
    pr.blast_bx = pr.blast_by = pr.blast_bz = 1.0 / np.sqrt(3.0)
    yield "blast3d_8", p, pr


def params_dict(p):
    d = {}
    for name, _ in A.Params._fields_:
        v = getattr(p, name)
        d[name] = list(v) if name == "bc" else v
    return d


def main():
    O.build()
    for name, p, pr in cases():
        s = O.Solver(p)
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
