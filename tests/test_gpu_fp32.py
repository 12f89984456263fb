"""FP32-bulk form on the GPU: the half step, the records and the face/edge solves in float;
the state, the update, constrained transport and the CFL reduction in double.

Bars (the error is stated against the FP64 oracle; there is no reference FP32 path):
per-step relative L-inf <= 1e-5 at fixed dt, div B at FP64 roundoff (CT is applied in double
to whatever EMFs arrive), mass/energy/momentum/face-field totals conserved to FP64 roundoff
(the double update telescopes the float fluxes exactly), and the 1000-step Orszag-Tang drift
from the FP64 solution bounded and reported.
"""
import numpy as np
import pytest

from parity import rel_l1, rel_linf
from ppkmhd_amd import _abi as A
from ppkmhd_amd import default_problem, make_params

pytestmark = pytest.mark.gpu


def solver(p, pr=None):
    from ppkmhd_amd.solver import SolverMHDMuscl
    return SolverMHDMuscl(p, pr)


def blast3d(n, **kw):
    p = make_params(3, n, n, n, gamma0=5 / 3, cfl=0.8, slope_type=1.0, tEnd=10.0, **kw)
    pr = default_problem(A.PROBLEM_BLAST, p)
    pr.blast_radius = 0.2
    pr.blast_density_out = 1.0
    pr.blast_bx = pr.blast_by = pr.blast_bz = 1.0 / np.sqrt(3.0)
    return p, pr


@pytest.mark.parametrize("dim", [2, 3])
def test_fp32_bulk_per_step(oracle, gpu, dim):
    if dim == 2:
        p = make_params(2, 64, 64, gamma0=5 / 3, cfl=0.8, tEnd=0.5, precision=A.FP32_BULK)
        pr = default_problem(A.PROBLEM_ORSZAG_TANG, p)
        dt = 2e-3
    else:
        p, pr = blast3d(16, precision=A.FP32_BULK)
        dt = 5e-3
    o = oracle.Solver(p)
    o.init_problem(pr)
    g = solver(p, pr)
    worst = 0.0
    for _ in range(5):
        g.upload(o.get_state())
        g.set_time(*o.time()[::2])
        o.step(dt, A.DT_FIXED)
        g.step(dt, A.DT_FIXED)
        worst = max(worst, rel_linf(g.download(), o.get_state(), p))
    print(f"FP32 bulk, {dim}D: worst relative L-inf per step {worst:.3e}")
    assert worst <= 1e-5, worst


def test_fp32_bulk_divb_and_conservation(gpu):
    p = make_params(3, 32, 32, 32, gamma0=5 / 3, cfl=0.8, tEnd=10.0, precision=A.FP32_BULK)
    pr = default_problem(A.PROBLEM_ORSZAG_TANG, p)
    pr.ot_kt = 1.0
    g = solver(p, pr)
    d0 = g.divb_max()
    s0 = g.conserved_sums()
    g.run(50)
    d1 = g.divb_max()
    s1 = g.conserved_sums()
    bscale = 1.0 / np.sqrt(4 * np.pi) / p.dx
    assert d0 < 1e-12 * bscale and d1 < 1e-12 * bscale, (d0, d1, bscale)
    np.testing.assert_allclose(s1, s0, rtol=1e-12, atol=1e-9)


def test_fp32_bulk_1000_steps_drift(oracle, gpu):
    """Orszag-Tang 2D 64^2, 1000 CFL steps: relative L1 distance to the FP64 oracle."""
    p = make_params(2, 64, 64, gamma0=5 / 3, cfl=0.8, tEnd=10.0, n_step_max=1000,
                    precision=A.FP32_BULK)
    pr = default_problem(A.PROBLEM_ORSZAG_TANG, p)
    o = oracle.Solver(p)
    o.init_problem(pr)
    g = solver(p, pr)
    for _ in range(1000):
        o.step(0, A.DT_CFL)
    g.run(1000)
    e = rel_l1(g.download(), o.get_state(), p)
    print(f"FP32 bulk, OT 2D 64^2 after 1000 steps: relative L1 {e:.3e}")
    assert e <= 1e-3, e
