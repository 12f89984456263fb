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


def test_solver_tables():
    """tests/golden/hlld_table.npz and emf_table.npz (written by make_golden_solvers.py from the
    oracle): the oracle reproduces every flux and EMF bit for bit, and the independent textbook
    HLLD (Miyoshi & Kusano 2005) and the branch form of the 2D HLLD EMF agree with the tables
    to 1e-11 relative."""
    import textbook as T
    from oracle import pyoracle as O
    from ppkmhd_amd import make_params

    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    h = np.load(os.path.join(here, "hlld_table.npz"))
    worst = 0.0
    for n in range(h["flux"].shape[0]):
        p = make_params(3, 8, 8, 8, gamma0=float(h["gamma"][n]), riemann=A.RIEMANN_HLLD,
                        smallc=1e-10, float_round=False)
        f = O.riemann(p, h["qL"][n], h["qR"][n])
        assert np.array_equal(f, h["flux"][n]), n
        g = T.hlld(h["qL"][n], h["qR"][n], p.gamma0)
        worst = max(worst, float(np.max(np.abs(g - f)) / max(1.0, float(np.max(np.abs(f))))))
    assert worst < 1e-11, worst
    e = np.load(os.path.join(here, "emf_table.npz"))
    p = make_params(3, 8, 8, 8, gamma0=float(e["gamma"]), riemann=A.RIEMANN_HLLD,
                    smallc=float(e["smallc"]), float_round=False)
    worst = 0.0
    for n in range(e["emf"].shape[0]):
        got = O.mag_riemann2d_hlld(p, e["q"][n], e["e"][n])
        assert got == e["emf"][n], n
        ref = T.mag_riemann2d_branchy(*e["q"][n], *e["e"][n], p.gamma0, p.smallc)
        worst = max(worst, abs(ref - got) / (1.0 + abs(got)))
    assert worst < 1e-11, worst
