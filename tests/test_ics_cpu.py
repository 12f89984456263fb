"""Host initial conditions of the library (ppk_mhd_initial_state, the state ppk_mhd_init_problem
uploads) against the oracle's, bit for bit, for every problem in 2D and 3D (no GPU)."""
import numpy as np
import pytest

from ppkmhd_amd import _abi as A
from ppkmhd_amd import default_problem, make_params
from ppkmhd_amd.solver import initial_state


def cases():
    out = []
    for dim in (2, 3):
        n = (12, 10, 8) if dim == 3 else (20, 14)
        for kind in (A.PROBLEM_ORSZAG_TANG, A.PROBLEM_BLAST, A.PROBLEM_KELVIN_HELMHOLTZ,
                     A.PROBLEM_WAVE, A.PROBLEM_IMPLODE):
            out.append((dim, n, kind, None))
        out.append((dim, n, A.PROBLEM_KELVIN_HELMHOLTZ, "rand"))
    out.append((3, (12, 10, 8), A.PROBLEM_KELVIN_HELMHOLTZ, "rand_y"))
    out.append((3, (12, 10, 8), A.PROBLEM_ORSZAG_TANG, "kt"))
    for t in (0, 1, 2):
        out.append((3, (12, 10, 8), A.PROBLEM_WAVE, f"wave{t}"))
    return out


@pytest.mark.parametrize("dim,n,kind,variant", cases())
def test_host_initial_state_equals_oracle(oracle, dim, n, kind, variant):
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
        pr.ot_b0_scale = 1.3
    elif variant and variant.startswith("wave"):
        pr.wave_type = int(variant[-1])
    U = initial_state(p, pr)
    ref = oracle.init_problem_array(p, pr)
    assert np.array_equal(U, ref)
    assert np.isfinite(U).all()
