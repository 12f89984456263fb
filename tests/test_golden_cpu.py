"""The oracle reproduces the committed golden fixtures bit for bit (regression pin, CPU)."""
import ast
import os

import numpy as np
import pytest

from ppkmhd_amd import _abi as A

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load_case(name):
    z = np.load(os.path.join(GOLDEN, name + ".npz"))
    pd = ast.literal_eval(str(z["params"]))
    p = A.Params()
    for k, v in pd.items():
        if k == "bc":
            for f in range(6):
                p.bc[f] = v[f]
        else:
            setattr(p, k, v)
    return p, z


@pytest.mark.parametrize("name", ["ot2d_16", "blast3d_8"])
def test_oracle_matches_golden(oracle, name):
    p, z = load_case(name)
    s = oracle.Solver(p)
    s.set_state(z["U0"])
    s.step(0, A.DT_CFL)
    assert np.array_equal(s.get_state(), z["U1"])
    assert s.time()[0] == z["time1"][0]
    for _ in range(9):
        s.step(0, A.DT_CFL)
    assert np.array_equal(s.get_state(), z["U10"])
    assert tuple(s.time()[:2]) == tuple(z["time10"])


@pytest.mark.parametrize("name", ["hydro_implode2d_32x16", "hydro_blast3d_16x12x8"])
def test_hydro_oracle_matches_golden(name):
    from oracle import pyhydro

    pyhydro.build()
    p, z = load_case(name)
    s = pyhydro.Solver(p)
    s.set_state(z["U0"])
    s.step(0, A.DT_CFL)
    assert np.array_equal(s.get_state(), z["U1"])
    assert s.time()[0] == z["time1"][0]
    for _ in range(9):
        s.step(0, A.DT_CFL)
    assert np.array_equal(s.get_state(), z["U10"])
    assert tuple(s.time()[:2]) == tuple(z["time10"])
