"""On-device initial conditions against the host ones (through the solver: initial state,
ghost fill, U2 = U, first dt).  Problems without transcendental functions per cell are equal
bit for bit; the others agree to the last bits of the device sin / cos / exp."""
import numpy as np
import pytest

from parity import rel_linf
from ppkmhd_amd import _abi as A
from ppkmhd_amd import default_problem, make_params
from test_ics_cpu import cases

pytestmark = pytest.mark.gpu

EXACT = (A.PROBLEM_BLAST, A.PROBLEM_IMPLODE, A.PROBLEM_ROTOR, A.PROBLEM_FIELD_LOOP)  # sqrt, / only


@pytest.mark.parametrize("dim,n,kind,variant",
                         [c for c in cases() if not (c[0] == 2 and c[2] == A.PROBLEM_WAVE)])
def test_device_initial_state_matches_host(gpu, dim, n, kind, variant):
    """(The 2D linear wave is an all-zero state in the reference: nothing to step.)"""
    from ppkmhd_amd.solver import SolverMHDMuscl
    p = make_params(dim, *n, xmin=-0.5, xmax=0.5, ymin=-0.5, ymax=0.5, zmin=-0.5, zmax=0.5,
                    gamma0=1.4)
    pr = default_problem(kind, p)
    if variant in ("rand", "rand_y"):
        pr.kh_p_rand, pr.kh_p_sine, pr.kh_p_sine_rob = 1, 0, 0
        pr.kh_seed = 77
        if variant == "rand_y":
            pr.kh_axis = 1
            pr.kh_bx = 0.1
    elif variant == "kt":
        pr.ot_kt = 1.0
    elif variant and variant.startswith("wave"):
        pr.wave_type = int(variant[-1])
    elif variant == "small":
        pr.rotor_r0, pr.rotor_r1 = 0.12, 0.3
        pr.loop_radius, pr.loop_density_in, pr.loop_vflow = 0.3, 2.5, 0.7
    a = SolverMHDMuscl(p)
    a.init_problem(pr)
    b = SolverMHDMuscl(p)
    b.init_problem(pr, on_device=True)
    Ua, Ub = a.download(), b.download()
    if kind in EXACT or variant in ("rand", "rand_y"):
        assert np.array_equal(Ua, Ub)
    else:
        assert rel_linf(Ub, Ua, p) <= 1e-14
    assert b.time()[1] == pytest.approx(a.time()[1], rel=1e-13)
    # and both evolve the same way
    a.run(3)
    b.run(3)
    assert rel_linf(b.download(), a.download(), p) <= 1e-12
