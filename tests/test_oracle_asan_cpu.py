"""The CPU oracles (test infrastructure) built for the host with AddressSanitizer and
UndefinedBehaviorSanitizer (tests/cxx/oracle_asan.cpp driving oracle/mhd_oracle.cpp, and
with -DORACLE_HYDRO oracle/hydro_oracle.cpp): a few CFL steps of several problems run with
every out-of-bounds access, use of freed memory, signed overflow, misaligned or invalid
load and division by zero an error, and the result equals the regular oracle library's bit
for bit (both builds keep -ffp-contract=off, so the optimisation level cannot change a
rounding).  Cases cover 2D and 3D, the MHD solvers HLLD / HLL / LLF, both slope limiters and
the three boundary types, and the hydro solvers on their own problems."""
import os
import subprocess

import numpy as np
import pytest

from ppkmhd_amd import _abi as A
from ppkmhd_amd import default_problem, make_params
from ppkmhd_amd.params import make_hydro_params

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "cxx", "oracle_asan.cpp")


def build(out, source, hydro):
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-ffp-contract=off",
           "-fno-fast-math", "-Wno-unknown-pragmas", "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "oracle", source), DRIVER, "-o", out]
    if hydro:
        cmd.insert(1, "-DORACLE_HYDRO")
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    return out


@pytest.fixture(scope="module")
def exes(tmp_path_factory):
    d = tmp_path_factory.mktemp("oracle_asan")
    return {"mhd": build(str(d / "mhd_asan"), "mhd_oracle.cpp", False),
            "hydro": build(str(d / "hydro_asan"), "hydro_oracle.cpp", True)}


def run(exe, tmp, p, pr, steps, shape):
    d = str(tmp)
    with open(os.path.join(d, "params.bin"), "wb") as f:
        f.write(bytes(p))
    with open(os.path.join(d, "problem.bin"), "wb") as f:
        f.write(bytes(pr))
    with open(os.path.join(d, "meta.txt"), "w") as f:
        f.write(f"{steps}\n")
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([exe, d], capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0, (r.stdout + r.stderr)[-4000:]
    U = np.fromfile(os.path.join(d, "u.bin"), dtype=np.float64).reshape(shape)
    t, dt, it = open(os.path.join(d, "time.txt")).read().split()
    return U, float(t), float(dt), int(it)


def mhd_cases():
    ot = make_params(2, 24, 20, gamma0=5 / 3, cfl=0.8, slope_type=1.0)
    yield "ot2d_hlld_minmod", ot, default_problem(A.PROBLEM_ORSZAG_TANG, ot), 6
    bl = make_params(3, 10, 9, 8, gamma0=5 / 3, cfl=0.8)
    pb = default_problem(A.PROBLEM_BLAST, bl)
    pb.blast_radius = 0.3
    yield "blast3d_hlld", bl, pb, 4
    kh = make_params(3, 9, 10, 7, gamma0=1.4, cfl=0.7, slope_type=2.0, riemann=A.RIEMANN_HLL,
                     bc=[A.BC_DIRICHLET, A.BC_NEUMANN, A.BC_PERIODIC, A.BC_PERIODIC,
                         A.BC_NEUMANN, A.BC_DIRICHLET])
    yield "kh3d_hll_mc_mixed", kh, default_problem(A.PROBLEM_KELVIN_HELMHOLTZ, kh), 3
    im = make_params(2, 18, 14, gamma0=1.4, cfl=0.6, riemann=A.RIEMANN_LLF,
                     bc=[A.BC_NEUMANN, A.BC_DIRICHLET, A.BC_DIRICHLET, A.BC_NEUMANN,
                         A.BC_PERIODIC, A.BC_PERIODIC])
    yield "implode2d_llf_walls", im, default_problem(A.PROBLEM_IMPLODE, im), 5


@pytest.mark.parametrize("case", list(mhd_cases()), ids=lambda c: c[0])
def test_mhd_oracle_under_sanitizers(exes, oracle, tmp_path, case):
    name, p, pr, steps = case
    o = oracle.Solver(p)
    o.init_problem(pr)
    for _ in range(steps):
        o.step(0.0, A.DT_CFL)
    ref = o.get_state()
    U, t, dt, it = run(exes["mhd"], tmp_path, p, pr, steps, ref.shape)
    assert np.isfinite(ref).all()
    assert np.array_equal(U, ref), f"{name}: sanitized oracle differs"
    assert (t, dt, it) == o.time()


def hydro_cases():
    im = make_hydro_params(2, 20, 16, gamma0=1.4, cfl=0.7, slope_type=2.0,
                           riemann=A.RIEMANN_HLLC, bc=[A.BC_DIRICHLET] * 6)
    yield "implode2d_hllc_walls", im, default_problem(A.PROBLEM_IMPLODE, im), 5
    bl = make_hydro_params(3, 9, 8, 10, gamma0=1.4, cfl=0.7, riemann=A.RIEMANN_APPROX,
                           bc=[A.BC_PERIODIC, A.BC_NEUMANN, A.BC_DIRICHLET, A.BC_PERIODIC,
                               A.BC_NEUMANN, A.BC_NEUMANN])
    yield "blast3d_approx_mixed", bl, default_problem(A.PROBLEM_BLAST, bl), 4


@pytest.mark.parametrize("case", list(hydro_cases()), ids=lambda c: c[0])
def test_hydro_oracle_under_sanitizers(exes, tmp_path, case):
    from oracle import pyhydro

    pyhydro.build()
    name, p, pr, steps = case
    o = pyhydro.Solver(p)
    o.init_problem(pr)
    for _ in range(steps):
        o.step(0.0, A.DT_CFL)
    ref = o.get_state()
    U, t, dt, it = run(exes["hydro"], tmp_path, p, pr, steps, ref.shape)
    assert np.isfinite(ref).all()
    assert np.array_equal(U, ref), f"{name}: sanitized hydro oracle differs"
    assert (t, dt, it) == o.time()
