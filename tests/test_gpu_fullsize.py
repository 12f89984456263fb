"""Parity and invariants at the benchmark sizes (GPU).

  blast 3D 256^3 (the headline workload): strict FP64 bit-identical to the oracle over 2 CFL
               steps, t and dt included; default FP64 within 1e-12 per step; FP32-bulk error
               per step stated
  Orszag-Tang 2D 256^2 to t = 0.5 (the CPU-runnable workload): default FP64 relative L1 to the
               oracle at most 1e-10 at t_end
  blast 3D 512^3: every value finite, conserved totals to 1e-12 and div B at roundoff after
               10 CFL steps (cell offsets above 2^31 doubles per variable plane set)
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


def blast(n, **kw):
    """The headline workload: rho 1, p 10 inside r < 0.1 and 0.1 outside, B = (1,1,1)/sqrt(3)."""
    p = make_params(3, n, n, n, gamma0=5 / 3, cfl=0.8, slope_type=1.0, tEnd=1.0,
                    n_step_max=200, **kw)
    pr = default_problem(A.PROBLEM_BLAST, p)
    pr.blast_radius = 0.1
    pr.blast_density_in = pr.blast_density_out = 1.0
    pr.blast_pressure_in, pr.blast_pressure_out = 10.0, 0.1
    pr.blast_bx = pr.blast_by = pr.blast_bz = 1.0 / np.sqrt(3.0)
    return p, pr


def test_blast_256_strict_bit_identical(oracle, gpu):
    p, pr = blast(256, precision=A.FP64_STRICT)
    o = oracle.Solver(p)
    o.init_problem(pr)
    g = solver(p, pr)
    for s in range(2):
        o.step(0, A.DT_CFL)
        g.step(0, A.DT_CFL)
        assert np.array_equal(g.download(), o.get_state()), f"step {s + 1} differs"
        assert g.time() == o.time()
    print("blast 256^3: strict FP64 bit-identical to the oracle over 2 CFL steps", flush=True)


@pytest.mark.parametrize("precision,bar", [(A.FP64, 1e-12), (A.FP32_BULK, 1e-5)])
def test_blast_256_per_step_error(oracle, gpu, precision, bar):
    p, pr = blast(256, precision=precision)
    o = oracle.Solver(p)
    o.init_problem(pr)
    g = solver(p, pr)
    worst = 0.0
    for _ in range(2):
        g.upload(o.get_state())
        g.set_time(*o.time()[::2])
        o.step(0, A.DT_CFL)
        g.step(0, A.DT_CFL)
        worst = max(worst, rel_linf(g.download(), o.get_state(), p))
    print(f"blast 256^3, precision {precision}: worst relative L-inf per step {worst:.3e}",
          flush=True)
    assert worst <= bar, worst


def test_orszag_tang_256_to_tend(oracle, gpu):
    p = make_params(2, 256, 256, gamma0=5 / 3, cfl=0.8, slope_type=1.0, tEnd=0.5,
                    n_step_max=100000)
    pr = default_problem(A.PROBLEM_ORSZAG_TANG, p)
    o = oracle.Solver(p)
    o.init_problem(pr)
    while not o.finished():
        o.step(0, A.DT_CFL)
    g = solver(p, pr)
    g.run(100000)
    assert g.finished() and g.time()[2] == o.time()[2]
    assert g.time()[0] == pytest.approx(o.time()[0], rel=1e-12)
    e = rel_l1(g.download(), o.get_state(), p)
    print(f"Orszag-Tang 256^2 to t = 0.5 ({o.time()[2]} steps): relative L1 {e:.3e}", flush=True)
    assert e <= 1e-10, e


def test_blast_512_invariants(gpu):
    p, pr = blast(512)
    g = solver(p, pr)
    s0 = g.conserved_sums()
    g.run(10)
    assert g.time()[2] == 10
    U = g.download()
    assert np.isfinite(U[p.interior()]).all()
    del U
    s1 = g.conserved_sums()
    scale = np.abs(s0).max()
    np.testing.assert_allclose(s1, s0, rtol=1e-12, atol=1e-12 * scale)
    d = g.divb_max()
    print(f"blast 512^3 after 10 steps: div B max {d:.2e}, totals {s1}", flush=True)
    assert d < 1e-12 / p.dx, d
